/*
 * dnsg.h -- C-ABI of the MI355X-native grid-embedded D3Q19 channel-flow time step.
 *
 * This header is the drop-in seam between the C driver (dnsg-run) and the HIP
 * kernels in libdnsg.so.  It replaces the per-step function table of the reference
 * solver (DNS-solver-with-grid-embedding/src/fun_defs.h:4-36), whose functions take
 * a POINTER state struct by value (definitions.h:161-229) and mutate its arrays.
 * Here the state lives on the GPU behind an opaque handle; host buffers are only
 * borrowed for the duration of a call and every function returns an int status.
 *
 * Axis convention (same as the reference): x streamwise (periodic), y spanwise
 * (periodic), z wall-normal.  Host-side array layouts are the reference's file
 * layouts without ghost cells:
 *   coarse lattice     [Xc][Yc][zdc][19]   (lbf-c.NNNN halves, main:1264-1274)
 *   fine band lattice  [Xf][Yf][zdf][19]   (lbf-slf / lbf-suf.NNNN, main:1276-1298)
 *   interface planes   [3][Xc][Yc][23]     (lbf-lin / lbf-uin.NNNN, main:1310-1340)
 * with Xc = xdv, Yc = ydv, Xf = GR*Xc, Yf = GR*Yc, zdf = zds + 1.
 *
 * No torch or HIP types appear in any signature: plain pointers and sizes only.
 */
#ifndef DNSG_H
#define DNSG_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNSG_ABI_VERSION 1

/* Wall surface pattern (definitions.h:70-72 PSLIP/RSLIP/SSLIP). */
enum dnsg_pattern {
    DNSG_PATTERN_NONE  = 0, /* all no-slip walls (solid_init.c:46-64 only)          */
    DNSG_PATTERN_RSLIP = 1, /* streamwise ridges   (slip_ridges_init.c:15-81)       */
    DNSG_PATTERN_PSLIP = 2, /* square posts        (slip_posts_init.c:15-132)       */
    DNSG_PATTERN_SSLIP = 3  /* transverse ridges   (slip_spanwise_init.c:15-80)     */
};

/* Body-force driving. */
enum dnsg_force_mode {
    DNSG_FORCE_FLUX  = 0, /* reference controller: G = ub - <rho u> (encalc2.c:129) */
    DNSG_FORCE_CONST = 1  /* constant body force g_const (coarse lattice units)      */
};

/* Storage / arithmetic tier. */
enum dnsg_precision {
    DNSG_PREC_FP64 = 0,   /* plain path: f stored and collided in FP64               */
    DNSG_PREC_FP32 = 1,   /* f - w*rho0 stored in FP32, FP32 collision               */
    DNSG_PREC_FP16 = 2    /* f - w*rho0 stored in FP16, FP32 collision               */
};

/* Initial state for dnsg_init_state. */
enum dnsg_init_kind {
    DNSG_INIT_REST          = 0, /* '-s' mode: f = ftemp = w on nodes with s > -1 (init.c:15-56) */
    DNSG_INIT_LAMINAR_NOISE = 1  /* rho = 1, u_x = u_max*4z(1-z) + seeded uniform noise, f = feq */
};

/* Status codes. */
enum dnsg_status {
    DNSG_OK        =  0,
    DNSG_ERR_ARG   = -1, /* invalid argument or parameter set                       */
    DNSG_ERR_HIP   = -2, /* HIP runtime error                                        */
    DNSG_ERR_NAN   = -3, /* NaN in the wall shear stress (wshearstr.c:139-143)       */
    DNSG_ERR_OOM   = -4, /* device allocation failed                                 */
    DNSG_ERR_STATE = -5, /* call out of sequence                                     */
    DNSG_ERR_IO    = -6  /* file read/write failed (driver-side helpers)             */
};

/*
 * Input parameters.  Field names mirror the compile-time #defines of the reference
 * (definitions.h:34-119); the trailing block holds runtime-only extensions that the
 * reference cannot express.  One process drives one GPU, so XDIM = YDIM = 1.
 */
typedef struct dnsg_params {
    /* definitions.h:34-41 */
    double Rebs;          /* bulk Reynolds number on 2h (tau formulas main:447-448) */
    double ub;            /* target bulk velocity (flux controller, encalc2.c:129)  */
    int    zdv;           /* coarse wall-normal points incl. the 2 overlap layers   */
    int    ydv;           /* coarse spanwise points                                  */
    int    xdv;           /* coarse streamwise points                                */
    int    TSTR;          /* starting time index (file numbering)                    */
    int    TEND;          /* end time index                                          */
    double CFL;           /* only defines utime (main:439-441)                       */
    /* definitions.h:47,51 */
    int    gratio;        /* GR: coarse/fine resolution ratio                        */
    int    zds;           /* fine fluid layers per band (zdf = zds + 1)              */
    /* definitions.h:58-62 (fine-cell units) */
    int    gyplussize, wyplussize, gxplussize, wxplussize;
    /* definitions.h:70-72 */
    int    pattern;       /* enum dnsg_pattern                                       */
    /* definitions.h:109 ITER: 0 = undefined (use (TEND-TSTR)*utime)                */
    long   ITER;
    /* ---- runtime-only extensions ---- */
    int    force_mode;    /* enum dnsg_force_mode                                    */
    double g_const;       /* body force for DNSG_FORCE_CONST, coarse lattice units   */
    double tau_c;         /* > 0: override coarse tau (tau_f = GR*(tau_c-0.5)+0.5)   */
    int    precision;     /* enum dnsg_precision                                     */
    int    init_kind;     /* enum dnsg_init_kind                                     */
    uint64_t seed;        /* counter-based PRNG seed for DNSG_INIT_LAMINAR_NOISE     */
    double u_max;         /* centreline velocity of the initial parabola             */
    double noise;         /* noise half-width as a fraction of u_max                 */
    /* definitions.h:87-90: 0 NOFLTR (ftcdtransfer.c), 1 FLTR (ftcdtransfer_lesfltr.c:
     * f_neq box-filtered over 5 x 5 x 5 fine nodes before injection; the reference's
     * filter arrays and weights assume gratio = 4, which is required) */
    int    fltr;
} dnsg_params;

/* Derived sizes and scalars (main:139-155, 439-469). */
typedef struct dnsg_dims {
    int    Xc, Yc, zdc;   /* coarse extents without ghosts                          */
    int    Xf, Yf, zdf;   /* fine band extents without ghosts (zdf includes the wall) */
    int    GR;
    int    NZs;           /* total wall-normal fine points incl. both wall nodes     */
    int    NZ;            /* statistics profile length 2*zdf + zdc - 4               */
    long   utime;         /* steps per unit non-dimensional time                     */
    long   up;            /* last step index of the time loop (loop runs 0..up)      */
    double dt;            /* time-stamp increment of the per-step logs               */
    double omega_c;       /* 1/tau_c (reference variable 'tauc', main:448-451)       */
    double omega_f;       /* 1/tau_f (reference variable 'tauf')                     */
    double tau_c, tau_f;
    size_t n_coarse;      /* Xc*Yc*zdc*19 doubles per coarse array                   */
    size_t n_band;        /* Xf*Yf*zdf*19 doubles per band array                     */
    size_t n_plane;       /* 3*Xc*Yc*23 doubles per interface-level file             */
} dnsg_dims;

/* Per-step scalars the driver appends to its text logs. */
typedef struct dnsg_diag {
    double time;          /* TM + dt*ss                                             */
    double force;         /* V.force[0] after encalc2 (coarse units)                */
    double momentum;      /* <rho u_x> weighted volume mean (encalc2.c:109)         */
    double density;       /* density.txt value (encalc2.c:142)                      */
    double flowrate;      /* flowrate.txt value (encalc2.c:138)                     */
    double str_up;        /* str-up.txt (wshearstr.c:134, 149)                      */
    double str_down;      /* str-do.txt (wshearstr.c:81, 154)                       */
    double stress;        /* w-str.txt  (wshearstr.c:138, 159)                      */
    double pgrad;         /* p-grad.txt (wshearstr.c:130-132, 164)                  */
} dnsg_diag;

typedef struct dnsg_ctx dnsg_t;

/* ---- parameters ---------------------------------------------------------- */

/* Fill *p with the defaults of definitions.h:34-119 scaled to one GPU. */
int dnsg_params_default(dnsg_params *p);

/* Derive sizes and relaxation rates (main:139-155, 432-469); validates the
 * pattern periods (the reference overflows when the extent is not a multiple of
 * the period, slip_ridges_init.c:40-42). */
int dnsg_dims_from_params(const dnsg_params *p, dnsg_dims *d);

/* ---- lifetime ------------------------------------------------------------- */

/* Allocate the device state (replaces the calloc block main:473-883).  The wall
 * masks are built on creation (solid_init + slip_*_init, main:964-973). */
int dnsg_create(const dnsg_params *p, dnsg_t **out);
int dnsg_destroy(dnsg_t *h);
const char *dnsg_last_error(const dnsg_t *h);
int dnsg_get_dims(const dnsg_t *h, dnsg_dims *d);

/* The node-code arrays V.s, V.sls and V.sus as solid_init (solid_init.c:21-133) and
 * the pattern routine (slip_ridges_init.c:15-81, slip_posts_init.c:15-132,
 * slip_spanwise_init.c:15-80) leave them on one rank (main:964-973), read back
 * from what this handle's kernels use: s_c [Xc+1][Yc+1][zdc], s_l and s_u
 * [Xf+1][Yf+1][zdf] (the ghosted extents of main:145-151), codes -1 no-slip wall,
 * -2 slip wall, 0 fluid, 1 periodic or bounded (solid_init.c:14-19).  Any pointer
 * may be NULL. */
int dnsg_node_codes(dnsg_t *h, int *s_c, int *s_l, int *s_u);

/* ---- state ---------------------------------------------------------------- */

/* '-s' mode init (init.c:15-56) or the seeded laminar+noise state, including the
 * three interface time levels sampled from it (xcomputations.c:558-625). */
int dnsg_init_state(dnsg_t *h);

/* '-p' mode read path (main:999-1183): upload the ftemp halves of lbf-c/slf/suf,
 * the three time levels of lbf-lin/uin (newest first) and force.NNNN.  Any array
 * pointer may be NULL to keep the current device contents. */
int dnsg_upload_state(dnsg_t *h, const double *coarse, const double *lower,
                      const double *upper, const double *lin3, const double *uin3,
                      double force);

/* The same state in the reference's in-memory arrays, so that the reference's main
 * can hand its own POINTER members over (V is passed by value, definitions.h:161-229):
 *   ftemp               V.ftemp    [Xc+1][Yc+1][zdc][19]  (main:145-146, 489-494)
 *   slftemp, suftemp    V.slftemp, V.suftemp [Xf+1][Yf+1][zdf][19]  (main:149-151)
 *   lin[3]              {V.lin, V.linp, V.linpp}, uin[3] {V.uin, V.uinp, V.uinpp},
 *                       each [Xc+6][Yc+6][23], interior at (i+2, j+2) (Fbi,
 *                       definitions.h:132; main:827-866)
 * The last x-plane and y-row of a lattice and the 2-cell ring (4 cells on the high
 * side) of a level are ghost cells: the upload reads only the interior, the download
 * writes every ghost cell as its periodic image, which is what the reference's halo
 * exchange puts there on one rank.  Any pointer may be NULL, as above. */
int dnsg_upload_state_ghosted(dnsg_t *h, const double *ftemp, const double *slftemp,
                              const double *suftemp, const double *const *lin,
                              const double *const *uin, double force);
int dnsg_download_state_ghosted(dnsg_t *h, double *ftemp, double *slftemp, double *suftemp,
                                double *const *lin, double *const *uin, double *force);

/* '-i' mode (main:982-997): replaces initializer() (initializer.c:15-545, the ftemp
 * pass, with sp_interpolate_qd.c) and inter_init() (inter_init.c:13-271).  `uniform`
 * is the ftemp half of the uniform-grid file lbf.NNNN, [xdu][ydu][zdu+2][19]
 * (definitions.h:107-109 sizes).  Every non-wall node of the three grids gets the
 * triquadratic interpolant of the uniform field, rescaled to its grid; the wall
 * layers start at rest; the interface levels follow inter_init, whose indexing
 * leaves the upper plane's rho and u in the lower level and copies only part of
 * each plane into the older levels.  The force starts at 0 (g_const in constant-
 * force mode).  DNSG_ERR_ARG if an interpolation stencil leaves the uniform array. */
int dnsg_init_uniform(dnsg_t *h, const double *uniform, int xdu, int ydu, int zdu);

/* Checkpoint write path (main:1262-1341): download the current state (the
 * reference's ftemp arrays) in the file layouts above. */
int dnsg_download_state(dnsg_t *h, double *coarse, double *lower, double *upper,
                        double *lin3, double *uin3, double *force);

/* ---- the time step ---------------------------------------------------------- */

/* Advance n coarse steps with the fused kernel sequence.  One step reproduces the
 * loop body main:1256-1375: encalc2, wshearstr, xcomputations, GR x
 * (xfcomputations, tinterpolate_hrmt, derivative, ctfdtransfer_opt, fexchange),
 * ftcdtransfer, exchange.  If diag != NULL it receives n records, one per step,
 * with the scalars encalc2/wshearstr compute at the top of that step.  Returns
 * DNSG_ERR_NAN (after filling the records up to the failing step) when the wall
 * stress is NaN. */
int dnsg_step(dnsg_t *h, int n, dnsg_diag *diag);

/* No counterpart in the reference: dnsg_step for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * dnsg_step may grow the handle's device record buffer and its pinned host staging
 * when n exceeds what an earlier call used; this variant allocates no device or host
 * memory on its success path.  The records of the n queued steps and the NaN halt
 * flag live in a fixed device array of the library (DNSG_NOALLOC_MAX_STEPS records,
 * about 288 KB, held by each GPU the library runs on until the process ends), and
 * come back through a buffer on the host stack.  All n steps are queued before the
 * one synchronisation, and the results, the records and the NaN behaviour are those
 * of dnsg_step.  n > DNSG_NOALLOC_MAX_STEPS returns DNSG_ERR_ARG.  Every handle on
 * one GPU shares that array, which keeps the last call's contents: calls that
 * overlap in time on one GPU (two host threads, two handles) race and must be
 * serialised by the caller. */
#define DNSG_NOALLOC_MAX_STEPS 4096
int dnsg_step_noalloc(dnsg_t *h, int n, dnsg_diag *diag);

/* Same as dnsg_step but without per-step host synchronisation; the diag records
 * are copied back once at the end.  *elapsed_ms receives the HIP-event time of
 * the n steps (device time, inputs resident in HBM). */
int dnsg_run_timed(dnsg_t *h, int n, dnsg_diag *diag, double *elapsed_ms);

/* dnsg_run_timed that also records HIP events in every timed step
 * (instrumentation, no counterpart in the reference) and returns in sec_ms[5] the
 * average device time per launch of the kernel groups of dnsg_kernel_times, in its
 * order, over those same steps.  group = -1 times every group (an event after each);
 * group = k in 0..4 times only group k (an event just before and after each of its
 * launches, the others 0), which adds fewer events to the timed steps.
 * sec_ms = NULL is dnsg_run_timed. */
int dnsg_run_timed_sections(dnsg_t *h, int n, dnsg_diag *diag, double *elapsed_ms, double *sec_ms, int group);

/* Instrumentation (no counterpart in the reference, which only times the whole
 * loop with MPI_Wtime, main:1252,1377): run n steps with HIP events between the
 * kernel groups on the handle's stream and return the average device time in ms
 * per launch of each group: ms[0] reductions (encalc2 + wshearstr),
 * ms[1] coarse collide-stream (with the bands' first substep unless
 * dnsg_band_launch_substeps gives 2), ms[2] interface sampling, ms[3] band
 * collide-stream (both bands: one launch per substep, or one launch for both
 * substeps where dnsg_band_launch_substeps gives 2), ms[4] interface fill
 * (+ fine->coarse injection after the last substep). */
int dnsg_kernel_times(dnsg_t *h, int n, double *ms);

/* Instrumentation (no counterpart in the reference): band substeps one band launch
 * of dnsg_step covers -- 2 where the handle runs both gratio-2 substeps in one pass
 * (the default at gratio 2 in FP64 and FP32 with constant force), else 1.  Read from the
 * environment at dnsg_create (no counterpart in the reference; every choice gives
 * the same state under constant force, and the same to the order of the reductions'
 * sums under the flux controller): DNSG_BAND2=1 / 0 forces the one-pass band kernel on / off at
 * gratio 2; DNSG_CS_PAIRED=0 keeps every collide-stream launch in its one-node form
 * (by default FP32 / FP16 rows of a multiple of 512 nodes, and FP64 ones in the
 * launches that sum their writes for the flux controller, take two nodes per
 * thread). */
int dnsg_band_launch_substeps(const dnsg_t *h, int *substeps);

/* Instrumentation (no counterpart in the reference): achievable HBM bandwidth of
 * a plain kernel over a buffer of `bytes` -- kind 0 copy (read + write),
 * 1 read-only, 2 write-only, 3 the collide-stream's data movement without its
 * arithmetic (FP64 lattice rows of 1024 x 512 x nz, 19 directions read and each
 * written into its (y, z) neighbour row; needs bytes >= 3 such layers) -- best of
 * `reps` launches, in GB/s of bytes read plus written. */
int dnsg_hbm_probe(size_t bytes, int kind, int reps, double *gbs);

/* ---- reference-granular entry points (one per fun_defs.h function) ---------- */

/* encalc2.c:17-147 -> V.force[0] (flux mode) and the flowrate/density scalars. */
int dnsg_encalc2(dnsg_t *h, long ss, dnsg_diag *d);
/* wshearstr.c:15-168 -> str-up/str-do/w-str/p-grad scalars; DNSG_ERR_NAN on NaN. */
int dnsg_wshearstr(dnsg_t *h, long ss, dnsg_diag *d);
/* xcomputations.c:16-950: coarse collide-stream, time-level rotation, sampling. */
int dnsg_xcomputations(dnsg_t *h);
/* xfcomputations.c:16-1144: one fine substep on both bands. */
int dnsg_xfcomputations(dnsg_t *h);
/* tinterpolate_hrmt.c + derivative.c + bcucof_opt.c + ctfdtransfer_opt.c for
 * substep iter (1..GR): fills the fine interface layers. */
int dnsg_ctfdtransfer(dnsg_t *h, int iter);
/* ftcdtransfer.c:16-118: fine -> coarse injection after the last substep. */
int dnsg_ftcdtransfer(dnsg_t *h);
/* exchange.c:16-25 / fexchange.c:15-28: swap current/next lattices. */
int dnsg_exchange(dnsg_t *h);
int dnsg_fexchange(dnsg_t *h);

/* ---- reductions and snapshots ----------------------------------------------- */

/* Global mass check (main:1193-1250): "Average Density in the channel". */
int dnsg_mass(dnsg_t *h, double *avg_density);

/* statistics.c:15-379: instantaneous plane statistics, 9*NZ doubles in the order
 * of turb-fld.000.NNN: U, V, W, u_rms, v_rms, w_rms, u'w', rho, rho_rms. */
int dnsg_statistics(dnsg_t *h, double *out);

/* Running (time-averaged) statistics accumulated on the device in FP64: folds
 * the current instantaneous profiles (as dnsg_statistics computes them) into the
 * accumulators.  The averages are plane-and-time means; second moments combine
 * the mean plane variance with the variance of the plane means. */
int dnsg_stats_accumulate(dnsg_t *h);
/* Returns the time averages: 9*NZ doubles as above (U, V, W, u_rms, v_rms, w_rms,
 * u'w', rho, rho_rms) plus *count samples. */
int dnsg_stats_average(dnsg_t *h, double *out, long *count);
int dnsg_stats_reset(dnsg_t *h);

/* vel_print.c:15-274: velocity and density of one grid; grid 0 = coarse,
 * 1 = lower band, 2 = upper band; vel is [X][Y][Z][3], den is [X][Y][Z]. */
int dnsg_velocity_snapshot(dnsg_t *h, int grid, double *vel, double *den);

/* e-spectra/src/1d-spectra-mgrd.c:119-230 (the reference's offline spectra tool, here
 * on the device from the current state): 1-D energy spectra of the velocity planes of
 * one grid, grid 0 = coarse, 1 = lower band, 2 = upper band.  Every wall-normal plane
 * of the grid's velocity file (vel_print.c; the bands include their wall layer) is
 * transformed over (x, y), scaled by 1/(N0 N1), its mean mode removed, and folded:
 *   ex [3][N0][nz]      E_uu, E_vv, E_ww over the streamwise wavenumber index
 *                        (the files Euux/Evvx/Ewwx-{c,fl,fu}.NNNN)
 *   ey [3][N1/2+1][nz]  the same over the spanwise wavenumber index (Euuy/Evvy/Ewwy)
 * with N0, N1 the grid's streamwise and spanwise extents, nz = zdc or zdf. */
int dnsg_spectra(dnsg_t *h, int grid, double *ex, double *ey);

/* TKE-budget/computations (budget2D.c:13-1736, uv_fluc.c:14-46; the reference's
 * offline budget tool, run here on the device from the current state): the
 * per-snapshot turbulence statistics and kinetic-energy budget profiles, NZ values
 * each (the profile layout of dnsg_statistics):
 *   turb [19][NZ]  UAVG VAVG WAVG UTIN VTIN WTIN uw DAVG DTIN SKEWu SKEWv SKEWw
 *                  SKEWuw SKEWp KURTu KURTv KURTw KURTuw KURTp (turb-fld-2d.000.NNNN)
 *   vort [3][NZ]   rms vorticity fluctuations wx wy wz       (vort-fluc-2d.000.NNNN)
 *   ke   [30][NZ]  Pii Puu Pvv Pww Puw Tiir Tuur Tvvr Twwr Tuwr Tiipai Tuupai Tvvpai
 *                  Twwpai Tuwpai Paiuu Paivv Paiww Paiuw Tiis Tuus Tvvs Twws Tuws
 *                  Epii Epuu Epvv Epww Epuw KE               (KE-Budget-2d.000.NNNN)
 * and, if ufl and vfl are both given, the lower band's u' and v' at k = 4, [Xf][Yf]
 * (ufl4/vfl4.NNNN).  The viscosity comes from Rebs and ub as in the tool.  The
 * layer sums are taken in another order than the tool's sequential loops.
 * Needs ydv >= 4 and zdv >= 6. */
int dnsg_budget2d(dnsg_t *h, double *turb, double *vort, double *ke, double *ufl, double *vfl);

/* Time averages of dnsg_budget2d's profiles over snapshots, in the accumulator
 * convention of the reference's serial post-processor (TKE-budget/postprocessing,
 * which averages the per-snapshot files: time_stat.c:132-210, vorticity.c:73-98,
 * Ebudget.c:184-300): the mean of each profile over the snapshots, except the rms
 * profiles -- turb rows UTIN VTIN WTIN DTIN and the three vort rows -- whose squares
 * are averaged (the tool takes the square root when it writes them).
 * dnsg_budget2d_accumulate runs dnsg_budget2d on the current state, folds its
 * profiles into the handle's FP64 sums and, where turb / vort / ke are not NULL,
 * returns them as dnsg_budget2d does (ufl / vfl are not formed).
 * dnsg_budget2d_average returns the averages and *count snapshots (DNSG_ERR_STATE
 * before the first fold); dnsg_budget2d_reset clears the sums. */
int dnsg_budget2d_accumulate(dnsg_t *h, double *turb, double *vort, double *ke);
int dnsg_budget2d_average(dnsg_t *h, double *turb, double *vort, double *ke, long *count);
int dnsg_budget2d_reset(dnsg_t *h);

/* ---- serving queue: steady slip-length solves, continuously batched -----------
 * No counterpart in the reference, which runs one channel per MPI job.  It is
 * for parametric sweeps: many independent channels of uneven sizes, each driven
 * by a constant body force (force_mode = DNSG_FORCE_CONST) until its bulk flux
 * is steady, each returning its effective slip length.
 *
 * Requests wait in arrival order and are admitted, first fit, while the
 * resident solves' device bytes stay within the budget and fewer than
 * max_active are resident.  All resident solves of one precision advance
 * together: every kernel of the step runs as one launch over up to 30 solves
 * (collide-stream, interface sampling, interface fill per gratio, injection,
 * flux sums).  Every check_every steps the bulk velocities of all resident
 * solves are read back in one copy; a solve whose bulk velocity changed by at
 * most tol (relative) over the last `window` steps retires, frees its memory
 * and makes room for the next admissions.  One host thread per queue. */
typedef struct dnsg_queue dnsg_queue_t;

#define DNSG_QUEUE_MAX_ACTIVE 256

typedef struct dnsg_queue_opts {
    size_t hbm_budget;  /* device bytes the resident solves may hold; 0 = free HBM at creation minus 4 GiB */
    int max_active;     /* resident solves at most, 1..DNSG_QUEUE_MAX_ACTIVE */
    int check_every;    /* steps between flux reads; divides window */
    int window;         /* convergence window in coarse steps */
    double tol;         /* converged when |U(t) - U(t - window)| <= tol |U(t)| */
    long max_steps;     /* a solve still moving after this many steps returns with status 1 */
    double floor_tol;   /* shifted tiers: steady at the storage precision when, after at least
                           5 windows, |U(t) - U(t - j window)| <= floor_tol |U(t)| for j = 1..5
                           and the 5 window-to-window changes have both signs (FP32 rounding
                           leaves U wandering near 1e-5 relative); 0 disables */
} dnsg_queue_opts;

typedef struct dnsg_slip_result {
    uint64_t id;
    int status;         /* 0 converged to tol, 2 steady within floor_tol (shifted tiers),
                           1 step cap reached, < 0 a dnsg error code */
    int precision;
    long steps;         /* coarse steps taken */
    double ubulk;       /* <rho u_x> / <rho> over the channel (encalc2.c:109 weights) */
    double u_noslip;    /* G H^2 / (12 nu_c), H = (NZs-2)/GR: the same channel with no-slip walls */
    double slip_length; /* b = (H/6)(ubulk/u_noslip - 1), coarse cells; equal slip on both walls */
    double t_arrival, t_admit, t_done; /* seconds on the queue clock */
} dnsg_slip_result;

/* window 1000, check_every 100, tol 1e-9, floor_tol 3e-5, max_steps 200000, max_active 64, budget 0 */
int dnsg_queue_opts_default(dnsg_queue_opts *o);
int dnsg_queue_create(const dnsg_queue_opts *o, dnsg_queue_t **out);
int dnsg_queue_destroy(dnsg_queue_t *q);
const char *dnsg_queue_last_error(const dnsg_queue_t *q);
/* seconds since the queue was created */
double dnsg_queue_clock(const dnsg_queue_t *q);
/* device bytes one solve takes while resident, as admission estimates it before
 * creating the solve: its buffers, each of 1 MiB or more rounded up to the runtime's
 * 2 MiB granularity, plus 2 MiB for the small ones (0 if the parameters are invalid).
 * Once admitted, the queue accounts the larger of this and the measured footprint
 * (free device memory before and after creating the solve). */
size_t dnsg_solve_bytes(const dnsg_params *p);
/* queue one solve; t_arrival < 0 stamps the current queue clock.  Rejects invalid
 * parameters, force_mode other than DNSG_FORCE_CONST, and a solve larger than
 * the whole budget, with DNSG_ERR_ARG. */
int dnsg_queue_submit(dnsg_queue_t *q, const dnsg_params *p, uint64_t id, double t_arrival);
/* Admit what fits, advance the resident solves by `steps` coarse steps (rounded
 * up to a multiple of check_every; returns early when nothing is resident or
 * waiting), and return up to cap finished solves in done (*ndone of them).
 * Finished solves beyond cap are kept for the next call. */
int dnsg_queue_advance(dnsg_queue_t *q, int steps, dnsg_slip_result *done, int cap, int *ndone);
int dnsg_queue_info(const dnsg_queue_t *q, int *pending, int *active, size_t *resident_bytes);
/* Check intervals run as a replayed HIP graph so far.  An interval (check_every
 * coarse steps of every resident solve) is captured once the resident set has stayed
 * the same for a whole interval and replayed while it stays so; the results are
 * those of the launches it records.  DNSG_QUEUE_GRAPH=0 in the environment at
 * dnsg_queue_create runs every interval with direct launches. */
int dnsg_queue_graph_launches(const dnsg_queue_t *q, long *launches);

#ifdef __cplusplus
}
#endif

#endif /* DNSG_H */
