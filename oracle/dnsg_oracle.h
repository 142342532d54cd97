/*
 * dnsg_oracle.h -- CPU restatement of the reference time step (TEST INFRASTRUCTURE).
 *
 * This is the oracle: a plain-C, single-process restatement of the reference
 * solver's time step (DNS-solver-with-grid-embedding/src), written in the
 * reference's own arithmetic order so that it reproduces the reference bit for
 * bit.  It exists to check the HIP path and to serve as the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may link,
 * load or execute it, and only as the checker or the CPU baseline -- never as the
 * product path.  The product (libdnsg.so) neither links nor loads it.
 *
 * Periodic form.  The reference carries a one-node ghost overlap in x and y that
 * its MPI halo exchange keeps equal to the periodic image (definitions.h:130-133,
 * main:145-151, xcomputations.c:206-546).  With one process every ghost is a copy
 * of node 0, so the oracle stores the lattices without ghosts and wraps x and y
 * indices instead.  State arrays use the reference's AoS file layouts
 * ([X][Y][Z][19] with q fastest), so checkpoints are byte-comparable with the
 * reference's lbf-* files.
 *
 * Pinning (how each function is checked):
 *   - the whole step, masks, logs, statistics, checkpoints: SHA-256 and text
 *     fixtures produced by the reference compiled from its own sources
 *     (oracle/build_ref.sh, tests/golden/make_golden.py);
 *   - Hermite, derivatives, bicubic, rescaling, collision invariants, streaming
 *     permutation, wall rules: closed-form tests in tests/test_oracle_math.py.
 */
#ifndef DNSG_ORACLE_H
#define DNSG_ORACLE_H

#include <stdint.h>
#include "../include/dnsg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* arrays reachable through orc_array() */
enum orc_array_id {
    ORC_C_F = 0, ORC_C_FT = 1,   /* coarse f / ftemp       [Xc][Yc][zdc][19] */
    ORC_L_F = 2, ORC_L_FT = 3,   /* lower band f / ftemp   [Xf][Yf][zdf][19] */
    ORC_U_F = 4, ORC_U_FT = 5,   /* upper band f / ftemp   [Xf][Yf][zdf][19] */
    ORC_LIN = 6, ORC_UIN = 7,    /* 3 time levels, newest first [3][Xc][Yc][23] */
    ORC_LINT = 8, ORC_UINT = 9   /* 4 slots (value, d/dx, d/dy, d2/dxdy) [4][Xc][Yc][23] */
};

orc_t *orc_create(const dnsg_params *p, char *err, int errlen);
void   orc_destroy(orc_t *o);
void   orc_get_dims(const orc_t *o, dnsg_dims *d);
void   orc_set_threads(int n);
double *orc_array(orc_t *o, int id);
double orc_get_force(const orc_t *o);
void   orc_set_force(orc_t *o, double g);

/* masks: periodic wall maps (-1 solid, -2 gas), [Xf][Yf] (x-major) */
void orc_wall_map(const orc_t *o, int upper, int8_t *out);
/* the reference's ghosted single-rank node-code arrays (solid_init.c + slip_*),
 * sizes (Xc+1)(Yc+1)zdc and (Xf+1)(Yf+1)zdf, [x][y][z] */
void orc_ref_masks(const orc_t *o, int *s_c, int *s_l, int *s_u);

/* initial states */
void orc_init_rest(orc_t *o);     /* init.c:15-56 ('-s' mode, walls zero)   */
void orc_init_seeded(orc_t *o);   /* seeded laminar profile + noise, f=feq  */
/* '-i' start (main:982-997, initializer.c + inter_init.c) from a uniform-grid lbf
 * field [2][xdu][ydu][zdu+2][19]; -1 with a message if its stencils leave the
 * uniform array; and a synthetic such field (seeded laminar profile + noise) */
int  orc_init_uniform(orc_t *o, const double *uni, int xdu, int ydu, int zdu, char *err, int errlen);
void orc_uniform_field(int xdu, int ydu, int zdu, uint64_t seed, double u_max, double noise, double *out);
uint64_t orc_prng_bits(uint64_t seed, int grid, int comp, uint64_t node);
double   orc_prng_sym(uint64_t seed, int grid, int comp, uint64_t node); /* [-1,1) */

/* the step, one reference function each */
void orc_encalc2(orc_t *o, long ss, dnsg_diag *d);
int  orc_wshearstr(orc_t *o, long ss, dnsg_diag *d);
void orc_xcomputations(orc_t *o);
void orc_xfcomputations(orc_t *o);
void orc_tinterpolate_hrmt(orc_t *o, int iter);
void orc_derivative(orc_t *o, double *addr);
void orc_bcucof(const orc_t *o, const double *addr, int ic, int jc, double *c);
void orc_ctfdtransfer(orc_t *o, int iter);
void orc_ftcdtransfer(orc_t *o);
void orc_exchange(orc_t *o);
void orc_fexchange(orc_t *o);
/* one full coarse step (main:1256-1375 minus file I/O); returns -3 on NaN stress */
int  orc_step(orc_t *o, long ss, dnsg_diag *d);
/* the same step with the body force replaced by g once encalc2 has formed it (tests) */
int  orc_step_forced(orc_t *o, long ss, dnsg_diag *d, double g);

/* reductions */
double orc_mass(const orc_t *o);                 /* main:1193-1250 */
void   orc_statistics(const orc_t *o, double *out); /* statistics.c:15-300, 9*NZ */

/* snapshots and profiles */
int  orc_vel_print(const orc_t *o, const char *dir, long cntr);   /* vel_print.c:15-274 */
int  orc_profile_len(const orc_t *o);                              /* 2*(zdf-GR-1)+zdc */
void orc_profile2d(const orc_t *o, double *jj, double *rho);       /* output_twoD2.c:62-124 */
void orc_get_levels(const orc_t *o, int upper, double *out);       /* [3][Xc][Yc][23], newest first */
void orc_set_levels(orc_t *o, int upper, const double *in);

/* single-node hooks (closed-form tests): lattice tables (e is [3][19]),
 * equilibrium and the BGK+force collision of xcomputations.c:59-85 */
void orc_tables(double *w, double *bi, int *e, int *opp, int *mir);
void orc_feq_node(double dens, const double *u, double *f);
void orc_collide_node(double *f, double G, double omega);

/* reference file layouts */
int orc_write_checkpoint(const orc_t *o, const char *dir, long index); /* main:1262-1341 */
int orc_read_checkpoint(orc_t *o, const char *dir, long index);        /* main:999-1183 */
/* append the per-step text logs (encalc2.c:135-144, wshearstr.c:145-166) */
int orc_append_logs(const char *dir, const dnsg_diag *d);
/* turb-fld.000.NNN (statistics.c:303-316) */
int orc_write_turbfld(const char *dir, long ts, const double *stats, int NZ);

/* run the reference driver loop (main:1254-1375) with its files, as a C API */
int orc_run(orc_t *o, const char *dir, long up, int write_files);

#ifdef __cplusplus
}
#endif
#endif
