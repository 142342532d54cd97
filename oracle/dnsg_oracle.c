/*
 * dnsg_oracle.c -- CPU restatement of the reference time step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
 * this file, as the checker or as the timed CPU baseline.  See dnsg_oracle.h for
 * the periodic single-process form and how each function is pinned.
 *
 * Every arithmetic expression follows the operation order of the reference file
 * and line cited next to it (no FMA contraction: build with -ffp-contract=off),
 * so results are bitwise equal to the reference's on one MPI rank.
 */
#define _POSIX_C_SOURCE 200809L
#include "dnsg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- lattice tables: globals.c:3-13 --------------------------------------- */
static const double w[19] = {1./3., 1./18., 1./18., 1./18., 1./18., 1./18., 1./18.,
                             1./36., 1./36., 1./36., 1./36., 1./36., 1./36.,
                             1./36., 1./36., 1./36., 1./36., 1./36., 1./36.};
static const double BI[19] = {0., 1./6., 1./6., 1./6., 1./6., 1./6., 1./6.,
                              1./12., 1./12., 1./12., 1./12., 1./12., 1./12.,
                              1./12., 1./12., 1./12., 1./12., 1./12., 1./12.};
static const int e0[19] = {0,1,-1,0,0,0,0,1,-1,-1,1,1,1,-1,-1,0,0,0,0};
static const int e1[19] = {0,0,0,1,-1,0,0,1,1,-1,-1,0,0,0,0,1,1,-1,-1};
static const int e2[19] = {0,0,0,0,0,1,-1,0,0,0,0,1,-1,-1,1,1,-1,-1,1};
static const double E0[19] = {0.,1.,-1.,0.,0.,0.,0.,1.,-1.,-1.,1.,1.,1.,-1.,-1.,0.,0.,0.,0.};
static const double E1[19] = {0.,0.,0.,1.,-1.,0.,0.,1.,1.,-1.,-1.,0.,0.,0.,0.,1.,1.,-1.,-1.};
static const double E2[19] = {0.,0.,0.,0.,0.,1.,-1.,0.,0.,0.,0.,1.,-1.,-1.,1.,1.,-1.,-1.,1.};
static const int opposite[19] = {0,2,1,4,3,6,5,9,10,7,8,13,14,11,12,17,18,15,16};
static const int mirror[19] = {0,1,2,3,4,6,5,7,8,9,10,12,11,14,13,16,15,18,17};

struct orc {
    dnsg_params p;
    dnsg_dims d;
    int Xc, Yc, zdc, Xf, Yf, zdf, GR, NZs, NZ;
    double tauc, tauf;             /* reference names: these hold omega = 1/tau  */
    double force[3], rho[3];       /* V.force, V.rho (definitions.h:163-164)      */
    double *cf, *cft;              /* V.f, V.ftemp                                */
    double *lf, *lft, *uf, *uft;   /* V.slf, V.slftemp, V.suf, V.suftemp          */
    double *lin, *linp, *linpp;    /* V.lin/linp/linpp, [Xc][Yc][23]              */
    double *uin, *uinp, *uinpp;
    double *lint, *uint_;          /* V.lint, V.uint: [4 slots][Xc][Yc][23]       */
    int8_t *wl, *wu;               /* wall codes of the lower/upper wall layers   */
};

#define CI(o, i, j, k) ((((size_t)(i) * (o)->Yc + (size_t)(j)) * (o)->zdc + (size_t)(k)) * 19)
#define FI(o, i, j, k) ((((size_t)(i) * (o)->Yf + (size_t)(j)) * (o)->zdf + (size_t)(k)) * 19)
#define PI(o, i, j) ((((size_t)(i)) * (o)->Yc + (size_t)(j)) * 23)
#define SLOT(o, s) ((size_t)(s) * (size_t)(o)->Xc * (size_t)(o)->Yc * 23)

static inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---- sizes and parameters: main:139-155, 432-469 ------------------------- */
static int derive(const dnsg_params *p, dnsg_dims *d, char *err, int errlen) {
    memset(d, 0, sizeof(*d));
    if (p->gratio < 1 || p->zds < 3 || p->zdv < 5 || p->xdv < 2 || p->ydv < 2) {
        snprintf(err, errlen, "bad sizes");
        return -1;
    }
    d->GR = p->gratio;
    d->Xc = p->xdv;          /* xdc = 1 + xdv/XDIM, unique nodes xdc-1 (main:145) */
    d->Yc = p->ydv;
    d->zdc = p->zdv;
    d->Xf = d->GR * d->Xc;   /* xdf = GR*(xdc-1)+1 (main:150) */
    d->Yf = d->GR * d->Yc;
    d->zdf = p->zds + 1;     /* main:152 */
    d->NZs = 2 * d->zdf + d->GR * (d->zdc - 4) + d->GR - 1; /* main:155 */
    d->NZ = 2 * d->zdf + (d->zdc - 4);                     /* statistics.c:26 */
    if (d->zdf - 1 - d->GR < 1 || d->zdf < 5) {
        snprintf(err, errlen, "zds too small for gratio");
        return -1;
    }
    {   /* main:439-441 */
        double ut = ((double)(0.5 * (d->NZs - 2.) / p->CFL)) / ((double)(d->GR));
        d->utime = (long)(int)ut;
        if (d->utime < 1) {
            snprintf(err, errlen, "CFL gives utime < 1");
            return -1;
        }
    }
    d->dt = d->GR * 1. / ((double)(0.5 * (d->NZs - 2.) / p->CFL)); /* encalc2.c:31 */
    if (p->tau_c > 0.) {
        d->tau_c = p->tau_c;
        d->tau_f = d->GR * (p->tau_c - 0.5) + 0.5;
    } else { /* main:447-448 */
        d->tau_f = (3. / p->Rebs) * p->ub * (d->NZs - 2.) + 0.5;
        d->tau_c = (1. / ((double)d->GR)) * (3. / p->Rebs) * p->ub * (d->NZs - 2.) + 0.5;
    }
    d->omega_f = 1. / d->tau_f; /* main:450-451 */
    d->omega_c = 1. / d->tau_c;
    {
        long nut = p->TEND - p->TSTR; /* main:464-465 */
        d->up = (long)(nut * (double)d->utime);
        if (p->ITER > 0) d->up = p->ITER; /* main:467-469 */
    }
    d->n_coarse = (size_t)d->Xc * d->Yc * d->zdc * 19;
    d->n_band = (size_t)d->Xf * d->Yf * d->zdf * 19;
    d->n_plane = (size_t)3 * d->Xc * d->Yc * 23;
    /* pattern extents must be whole periods (the reference overflows otherwise,
       slip_ridges_init.c:40-42, slip_posts_init.c:54-60) */
    if (p->pattern == DNSG_PATTERN_RSLIP || p->pattern == DNSG_PATTERN_PSLIP) {
        int P = p->gyplussize + p->wyplussize;
        if (P <= 0 || d->Yf % P) {
            snprintf(err, errlen, "spanwise pattern period must divide Yf");
            return -1;
        }
    }
    if (p->pattern == DNSG_PATTERN_SSLIP || p->pattern == DNSG_PATTERN_PSLIP) {
        int P = p->gxplussize + p->wxplussize;
        if (P <= 0 || d->Xf % P) {
            snprintf(err, errlen, "streamwise pattern period must divide Xf");
            return -1;
        }
    }
    return 0;
}

/* ---- masks: solid_init.c:21-133, slip_ridges_init.c:15-81,
 *      slip_posts_init.c:15-132, slip_spanwise_init.c:15-80 ------------------ */
static void stripe_map(int n, int gap, int width, int *g) {
    int j, a;
    for (j = 0; j < n; j++) g[j] = -1;                   /* slip_ridges_init.c:37-38 */
    for (j = width / 2; j < n; j += (gap + width))       /* :40-42 */
        for (a = 0; a < gap && j + a < n; a++) g[j + a] = -2;
}

static void build_walls(orc_t *o) {
    const dnsg_params *p = &o->p;
    int *sx = malloc(sizeof(int) * o->Xf), *sy = malloc(sizeof(int) * o->Yf);
    int i, j;
    for (i = 0; i < o->Xf; i++) sx[i] = -1;
    for (j = 0; j < o->Yf; j++) sy[j] = -1;
    if (p->pattern == DNSG_PATTERN_RSLIP) stripe_map(o->Yf, p->gyplussize, p->wyplussize, sy);
    if (p->pattern == DNSG_PATTERN_SSLIP) stripe_map(o->Xf, p->gxplussize, p->wxplussize, sx);
    if (p->pattern == DNSG_PATTERN_PSLIP) {
        stripe_map(o->Xf, p->gxplussize, p->wxplussize, sx);
        stripe_map(o->Yf, p->gyplussize, p->wyplussize, sy);
    }
    for (i = 0; i < o->Xf; i++)
        for (j = 0; j < o->Yf; j++) {
            int c = -1;
            if (p->pattern == DNSG_PATTERN_RSLIP) c = sy[j];
            else if (p->pattern == DNSG_PATTERN_SSLIP) c = sx[i];
            else if (p->pattern == DNSG_PATTERN_PSLIP) c = (sy[j] == -2) ? -2 : sx[i];
            o->wl[(size_t)i * o->Yf + j] = (int8_t)c;
            o->wu[(size_t)i * o->Yf + j] = (int8_t)c;
        }
    free(sx);
    free(sy);
}

void orc_wall_map(const orc_t *o, int upper, int8_t *out) {
    memcpy(out, upper ? o->wu : o->wl, (size_t)o->Xf * o->Yf);
}

void orc_ref_masks(const orc_t *o, int *s_c, int *s_l, int *s_u) {
    /* solid_init.c:21-133 on one rank (xdc = Xc+1, xdf = Xf+1), then the pattern
       maps with the ghost row/column taken from the periodic image */
    int xdc = o->Xc + 1, ydc = o->Yc + 1, zdc = o->zdc;
    int xdf = o->Xf + 1, ydf = o->Yf + 1, zdf = o->zdf;
    int i, j, k;
#define SC(i, j, k) s_c[((size_t)(i) * ydc + (j)) * zdc + (k)]
#define SL(i, j, k) s_l[((size_t)(i) * ydf + (j)) * zdf + (k)]
#define SU(i, j, k) s_u[((size_t)(i) * ydf + (j)) * zdf + (k)]
    for (i = 0; i < xdc; i++)
        for (j = 0; j < ydc; j++)
            for (k = 0; k < zdc; k++) {
                int edge = (k == 0 || k == zdc - 1 || i == 0 || i == xdc - 1 || j == 0 || j == ydc - 1);
                SC(i, j, k) = edge ? 1 : 0;
            }
    for (i = 0; i < xdf; i++)
        for (j = 0; j < ydf; j++)
            for (k = 0; k < zdf; k++) {
                int edge = (i == 0 || i == xdf - 1 || j == 0 || j == ydf - 1);
                int wc = o->wl[(size_t)(i % o->Xf) * o->Yf + (j % o->Yf)];
                /* lower band: k=0 wall, k=zdf-1 bounded, x/y edges bounded for k>=1 */
                SL(i, j, k) = (k == 0) ? wc : ((k == zdf - 1 || edge) ? 1 : 0);
                wc = o->wu[(size_t)(i % o->Xf) * o->Yf + (j % o->Yf)];
                /* upper band: k=zdf-1 wall, k=0 bounded, x/y edges bounded for k<zdf-1 */
                SU(i, j, k) = (k == zdf - 1) ? wc : ((k == 0 || edge) ? 1 : 0);
            }
#undef SC
#undef SL
#undef SU
}

/* ---- creation ---------------------------------------------------------------- */
orc_t *orc_create(const dnsg_params *p, char *err, int errlen) {
    char dummy[8];
    orc_t *o;
    if (!err) { err = dummy; errlen = sizeof(dummy); }
    o = calloc(1, sizeof(*o));
    if (!o) return NULL;
    o->p = *p;
    if (derive(p, &o->d, err, errlen)) { free(o); return NULL; }
    o->Xc = o->d.Xc; o->Yc = o->d.Yc; o->zdc = o->d.zdc;
    o->Xf = o->d.Xf; o->Yf = o->d.Yf; o->zdf = o->d.zdf;
    o->GR = o->d.GR; o->NZs = o->d.NZs; o->NZ = o->d.NZ;
    o->tauc = o->d.omega_c;
    o->tauf = o->d.omega_f;
    o->cf = calloc(o->d.n_coarse, 8);  o->cft = calloc(o->d.n_coarse, 8);
    o->lf = calloc(o->d.n_band, 8);    o->lft = calloc(o->d.n_band, 8);
    o->uf = calloc(o->d.n_band, 8);    o->uft = calloc(o->d.n_band, 8);
    {
        size_t np = (size_t)o->Xc * o->Yc * 23;
        o->lin = calloc(np, 8); o->linp = calloc(np, 8); o->linpp = calloc(np, 8);
        o->uin = calloc(np, 8); o->uinp = calloc(np, 8); o->uinpp = calloc(np, 8);
        o->lint = calloc(4 * np, 8); o->uint_ = calloc(4 * np, 8);
    }
    o->wl = calloc((size_t)o->Xf * o->Yf, 1);
    o->wu = calloc((size_t)o->Xf * o->Yf, 1);
    if (!o->cf || !o->cft || !o->lf || !o->lft || !o->uf || !o->uft || !o->lin || !o->linp ||
        !o->linpp || !o->uin || !o->uinp || !o->uinpp || !o->lint || !o->uint_ || !o->wl || !o->wu) {
        snprintf(err, errlen, "out of memory");
        orc_destroy(o);
        return NULL;
    }
    build_walls(o);
    o->force[0] = (p->force_mode == DNSG_FORCE_CONST) ? p->g_const : 0.; /* main:454 */
    return o;
}

void orc_destroy(orc_t *o) {
    if (!o) return;
    free(o->cf); free(o->cft); free(o->lf); free(o->lft); free(o->uf); free(o->uft);
    free(o->lin); free(o->linp); free(o->linpp); free(o->uin); free(o->uinp); free(o->uinpp);
    free(o->lint); free(o->uint_); free(o->wl); free(o->wu);
    free(o);
}

void orc_get_dims(const orc_t *o, dnsg_dims *d) { *d = o->d; }
double orc_get_force(const orc_t *o) { return o->force[0]; }
void orc_set_force(orc_t *o, double g) { o->force[0] = g; }

double *orc_array(orc_t *o, int id) {
    switch (id) {
    case ORC_C_F: return o->cf;
    case ORC_C_FT: return o->cft;
    case ORC_L_F: return o->lf;
    case ORC_L_FT: return o->lft;
    case ORC_U_F: return o->uf;
    case ORC_U_FT: return o->uft;
    case ORC_LIN: return o->lin;
    case ORC_UIN: return o->uin;
    case ORC_LINT: return o->lint;
    case ORC_UINT: return o->uint_;
    default: return NULL;
    }
}

void orc_get_levels(const orc_t *o, int upper, double *out) {
    size_t np = (size_t)o->Xc * o->Yc * 23;
    memcpy(out, upper ? o->uin : o->lin, np * 8);
    memcpy(out + np, upper ? o->uinp : o->linp, np * 8);
    memcpy(out + 2 * np, upper ? o->uinpp : o->linpp, np * 8);
}

void orc_set_levels(orc_t *o, int upper, const double *in) {
    size_t np = (size_t)o->Xc * o->Yc * 23;
    memcpy(upper ? o->uin : o->lin, in, np * 8);
    memcpy(upper ? o->uinp : o->linp, in + np, np * 8);
    memcpy(upper ? o->uinpp : o->linpp, in + 2 * np, np * 8);
}

/* ---- moments, equilibrium, collision (xcomputations.c:59-85) --------------- */
static inline void moments(const double *f, double *dens, double *u) {
    double ux = 0., uy = 0., uz = 0., d = 0.;
    int a;
    for (a = 0; a < 19; a++) {
        ux += E0[a] * f[a];
        uy += E1[a] * f[a];
        uz += E2[a] * f[a];
        d += f[a];
    }
    *dens = d;
    u[0] = ux / d;
    u[1] = uy / d;
    u[2] = uz / d;
}

static inline double feq_of(int a, double dens, const double *ueq, double v) {
    double u = E0[a] * ueq[0] + E1[a] * ueq[1] + E2[a] * ueq[2];
    return w[a] * dens * (1. + 3. * u + 4.5 * u * u - 1.5 * v);
}

static inline void collide(double *f, double Gx, double tau) {
    double dens, ueq[3], v;
    int a;
    moments(f, &dens, ueq);
    v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
    for (a = 0; a < 19; a++) {
        double feq = feq_of(a, dens, ueq, v);
        double Fi = BI[a] * dens * (E0[a] * Gx);
        f[a] += (Fi - tau * (f[a] - feq));
    }
}


/* ---- single-node hooks for the closed-form tests ---------------------------------- */
void orc_tables(double *wq, double *bi, int *e, int *opp, int *mir) {
    int a;
    for (a = 0; a < 19; a++) {
        wq[a] = w[a]; bi[a] = BI[a];
        e[a] = e0[a]; e[19 + a] = e1[a]; e[38 + a] = e2[a];
        opp[a] = opposite[a]; mir[a] = mirror[a];
    }
}
void orc_feq_node(double dens, const double *u, double *f) {
    double v = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    int a;
    for (a = 0; a < 19; a++) f[a] = feq_of(a, dens, u, v);
}
void orc_collide_node(double *f, double G, double omega) { collide(f, G, omega); }

/* ---- initial states ----------------------------------------------------------- */
void orc_init_rest(orc_t *o) { /* init.c:15-56; walls stay zero (calloc) */
    size_t n;
    int i, j, k, a;
    memset(o->cft, 0, o->d.n_coarse * 8); memset(o->cf, 0, o->d.n_coarse * 8);
    memset(o->lft, 0, o->d.n_band * 8);   memset(o->lf, 0, o->d.n_band * 8);
    memset(o->uft, 0, o->d.n_band * 8);   memset(o->uf, 0, o->d.n_band * 8);
    for (n = 0; n < o->d.n_coarse; n += 19)
        for (a = 0; a < 19; a++) { o->cf[n + a] = w[a]; o->cft[n + a] = w[a]; }
    for (i = 0; i < o->Xf; i++)
        for (j = 0; j < o->Yf; j++)
            for (k = 0; k < o->zdf; k++)
                for (a = 0; a < 19; a++) {
                    if (k != 0) { o->lf[FI(o, i, j, k) + a] = w[a]; o->lft[FI(o, i, j, k) + a] = w[a]; }
                    if (k != o->zdf - 1) { o->uf[FI(o, i, j, k) + a] = w[a]; o->uft[FI(o, i, j, k) + a] = w[a]; }
                }
    /* interface levels start at zero as in the calloc'd reference arrays */
    memset(o->lin, 0, (size_t)o->Xc * o->Yc * 23 * 8);
    memset(o->linp, 0, (size_t)o->Xc * o->Yc * 23 * 8);
    memset(o->linpp, 0, (size_t)o->Xc * o->Yc * 23 * 8);
    memset(o->uin, 0, (size_t)o->Xc * o->Yc * 23 * 8);
    memset(o->uinp, 0, (size_t)o->Xc * o->Yc * 23 * 8);
    memset(o->uinpp, 0, (size_t)o->Xc * o->Yc * 23 * 8);
}

/* counter-based PRNG: splitmix64 finaliser over (seed, grid, component, node) */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t orc_prng_bits(uint64_t seed, int grid, int comp, uint64_t node) {
    return mix64(mix64(seed) ^ mix64(((uint64_t)(grid & 0xff) << 56) ^ ((uint64_t)(comp & 0xff) << 48) ^ node));
}
double orc_prng_sym(uint64_t seed, int grid, int comp, uint64_t node) {
    double u = (double)(orc_prng_bits(seed, grid, comp, node) >> 11) * 0x1.0p-53;
    return 2. * u - 1.;
}

/* equilibrium of (rho=1, parabola + noise) at wall distance dist (fine units) */
static void seeded_node(const orc_t *o, int grid, uint64_t node, double dist, double *f) {
    double zeta = dist / ((double)(o->NZs - 2));
    double amp = o->p.noise * o->p.u_max;
    double ueq[3], v, dens = 1.;
    int a;
    ueq[0] = 4. * o->p.u_max * zeta * (1. - zeta) + amp * orc_prng_sym(o->p.seed, grid, 0, node);
    ueq[1] = amp * orc_prng_sym(o->p.seed, grid, 1, node);
    ueq[2] = amp * orc_prng_sym(o->p.seed, grid, 2, node);
    v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
    for (a = 0; a < 19; a++) f[a] = feq_of(a, dens, ueq, v);
}

static void sample_plane(const orc_t *o, const double *f, int k, double *plane);

void orc_init_seeded(orc_t *o) {
    int i, j, k, a;
    /* coarse: wall distance GR*(k-1)+(zdf-1-0.5) (statistics.c:322-323) */
#pragma omp parallel for private(j, k, a) schedule(static)
    for (i = 0; i < o->Xc; i++)
        for (j = 0; j < o->Yc; j++)
            for (k = 0; k < o->zdc; k++) {
                size_t n = CI(o, i, j, k);
                seeded_node(o, 0, n / 19, o->GR * (k - 1) + (o->zdf - 1 - 0.5), o->cft + n);
                for (a = 0; a < 19; a++) o->cf[n + a] = o->cft[n + a];
            }
#pragma omp parallel for private(j, k, a) schedule(static)
    for (i = 0; i < o->Xf; i++)
        for (j = 0; j < o->Yf; j++)
            for (k = 0; k < o->zdf; k++) {
                size_t n = FI(o, i, j, k);
                /* lower band: distance k-0.5 (statistics.c:320-321); wall node at rest */
                if (k == 0) for (a = 0; a < 19; a++) o->lft[n + a] = w[a];
                else seeded_node(o, 1, n / 19, k - 0.5, o->lft + n);
                /* upper band: distance k + GR*(zdc-3) + zdf-1-0.5 (statistics.c:324-325) */
                if (k == o->zdf - 1) for (a = 0; a < 19; a++) o->uft[n + a] = w[a];
                else seeded_node(o, 2, n / 19, k + o->GR * (o->zdc - 3) + o->zdf - 1. - 0.5, o->uft + n);
                for (a = 0; a < 19; a++) { o->lf[n + a] = o->lft[n + a]; o->uf[n + a] = o->uft[n + a]; }
            }
    /* three equal time levels sampled as xcomputations.c:558-625 does */
    sample_plane(o, o->cft, 1, o->lin);
    sample_plane(o, o->cft, o->zdc - 2, o->uin);
    memcpy(o->linp, o->lin, (size_t)o->Xc * o->Yc * 23 * 8);
    memcpy(o->linpp, o->lin, (size_t)o->Xc * o->Yc * 23 * 8);
    memcpy(o->uinp, o->uin, (size_t)o->Xc * o->Yc * 23 * 8);
    memcpy(o->uinpp, o->uin, (size_t)o->Xc * o->Yc * 23 * 8);
}

/* ---- '-i' start: initializer.c:15-819 (with sp_interpolate_qd.c:22-45), then
 * inter_init.c:13-271; main:982-997 ---------------------------------------------
 * The uniform field is the uniform-grid solver's lbf file, [2][xdu][ydu][zdu+2][19]
 * (ftemp half, then f half).  With one rank its x/y halo (2 nodes below, 4 above,
 * filled by the exchange of initializer.c:306-318) holds periodic images, so it is
 * read with wrapped x/y indices.  Every non-wall node of the three grids gets the
 * triquadratic Lagrange interpolant of the 3 x 3 x 3 uniform nodes around it,
 * rescaled f = feq + (dx_u/dx_g)(omega_u/omega_g) f_neq. */
typedef struct {
    const double *f;
    int xu, yu, zu; /* zu = zdu + 2 */
} uni_t;

static inline int wrapm(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

/* sp_interpolate_qd.c:22-45: Lagrange weights on (-1, 0, 1), z then y then x */
static void interp_qd(const uni_t *u, const int *ind, const int *jnd, const int *knd, double dx, double dy,
                      double dz, double *ans) {
    double cz[3][3][19], cy[3][19];
    int i, j, a;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) {
            const size_t col = ((size_t)wrapm(ind[i], u->xu) * u->yu + (size_t)wrapm(jnd[j], u->yu)) * u->zu;
            const double *f0 = u->f + (col + knd[0]) * 19, *f1 = u->f + (col + knd[1]) * 19,
                         *f2 = u->f + (col + knd[2]) * 19;
            for (a = 0; a < 19; a++)
                cz[i][j][a] = 0.5 * dz * (dz - 1.) * f0[a] - (dz - 1.) * (dz + 1.) * f1[a] + 0.5 * (dz + 1.) * dz * f2[a];
        }
    for (i = 0; i < 3; i++)
        for (a = 0; a < 19; a++)
            cy[i][a] = 0.5 * dy * (dy - 1.) * cz[i][0][a] - (dy - 1.) * (dy + 1.) * cz[i][1][a] +
                       0.5 * (dy + 1.) * dy * cz[i][2][a];
    for (a = 0; a < 19; a++)
        ans[a] = 0.5 * dx * (dx - 1.) * cy[0][a] - (dx - 1.) * (dx + 1.) * cy[1][a] + 0.5 * (dx + 1.) * dx * cy[2][a];
}

/* the 3 uniform indices per axis and the offsets (in uniform spacings) of node
 * (i, j, k) of grid g (0 coarse, 1 lower band, 2 upper band).  half 1 is the second
 * pass (initializer.c:567-819), which writes two expressions differently from the
 * first: the lower band's x/y reference index groups (i+0.5)*dxf/dxr otherwise
 * (:331-332 against :570-571) and the upper band's z offset is formed as
 * z + dxr/2 - kref dxr (:494 against :731). */
static void uni_stencil(const orc_t *o, double dxf, double dxc, double dxr, int g, int half, int i, int j, int k,
                        int *ind, int *jnd, int *knd, double *dx, double *dy, double *dz) {
    const int GR = o->GR, zdf = o->zdf, zdc = o->zdc;
    int iref, jref, kref, t;
    if (g == 1) { /* :326-366, :567-607 */
        if (half == 0) {
            iref = (int)floor((i + 0.5) * (dxf / dxr) - 0.5);
            jref = (int)floor((j + 0.5) * (dxf / dxr) - 0.5);
        } else {
            iref = (int)floor((i + 0.5) * dxf / dxr - 0.5);
            jref = (int)floor((j + 0.5) * dxf / dxr - 0.5);
        }
        kref = (int)floor((k - 0.5) * (dxf / dxr) + 0.5);
        *dx = (i + 0.5) * dxf - (iref + 0.5) * dxr;
        *dy = (j + 0.5) * dxf - (jref + 0.5) * dxr;
        if (kref == 0) { /* one-sided at the wall */
            for (t = 0; t < 3; t++) knd[t] = kref + t;
            *dz = -dxr + (k - 0.5) * dxf + 0.5 * dxr - kref * dxr;
        } else {
            for (t = 0; t < 3; t++) knd[t] = kref - 1 + t;
            *dz = (k - 0.5) * dxf + 0.5 * dxr - kref * dxr;
        }
    } else if (g == 0) { /* :403-431, :642-670 */
        iref = (int)floor((i + 0.5) * dxc / dxr - 0.5);
        jref = (int)floor((j + 0.5) * dxc / dxr - 0.5);
        kref = (int)floor(0.5 + (k * GR + (zdf - GR - 1 - 0.5)) * dxf / dxr);
        *dx = (i + 0.5) * dxc - (iref + 0.5) * dxr;
        *dy = (j + 0.5) * dxc - (jref + 0.5) * dxr;
        for (t = 0; t < 3; t++) knd[t] = kref - 1 + t;
        *dz = (k * GR + (zdf - GR - 1 - 0.5)) * dxf - (kref - 0.5) * dxr;
    } else { /* :470-499, :706-734 */
        iref = (int)floor((i + 0.5) * dxf / dxr - 0.5);
        jref = (int)floor((j + 0.5) * dxf / dxr - 0.5);
        kref = (int)floor(0.5 + (k + GR * (zdc - 3) + zdf - 1 - 0.5) * dxf / dxr);
        *dx = (i + 0.5) * dxf - (iref + 0.5) * dxr;
        *dy = (j + 0.5) * dxf - (jref + 0.5) * dxr;
        for (t = 0; t < 3; t++) knd[t] = kref - 1 + t;
        if (half == 0) *dz = (k + GR * (zdc - 3) + zdf - 1 - 0.5) * dxf - (kref - 0.5) * dxr;
        else *dz = (k + GR * (zdc - 3) + zdf - 1 - 0.5) * dxf + 0.5 * dxr - kref * dxr; /* :731 */
    }
    for (t = 0; t < 3; t++) {
        ind[t] = iref - 1 + t;
        jnd[t] = jref - 1 + t;
    }
    *dx = *dx / dxr;
    *dy = *dy / dxr;
    *dz = *dz / dxr;
}

/* f = feq + fac (f_int - feq) from the interpolated populations (:367-392) */
static void uni_rescale(const double *fi, double fac, double *out) {
    double dens, ueq[3], v;
    int a;
    moments(fi, &dens, ueq);
    v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
    for (a = 0; a < 19; a++) {
        const double feq = feq_of(a, dens, ueq, v);
        out[a] = feq + fac * (fi[a] - feq);
    }
}

/* the node ranges of the three destination grids: x, y extents and k in [klo, khi) */
static void uni_ranges(const orc_t *o, int g, int *nx, int *ny, int *klo, int *khi) {
    *nx = g == 0 ? o->Xc : o->Xf;
    *ny = g == 0 ? o->Yc : o->Yf;
    *klo = g == 1 ? 1 : 0;                       /* lower band: k = 1 .. zdf-1      */
    *khi = g == 0 ? o->zdc : (g == 1 ? o->zdf : o->zdf - 1); /* upper: k = 0 .. zdf-2 */
}

int orc_init_uniform(orc_t *o, const double *uni, int xdu, int ydu, int zdu, char *err, int errlen) {
    const int zu = zdu + 2; /* main:140 */
    const double dxf = 2. / ((double)(o->NZs - 2.)), dxc = o->GR * dxf, dxr = 2. / ((double)(zu - 2.));
    /* main:443-444: the uniform grid's omega from the same Reynolds number and bulk velocity */
    const double omega_u = 1. / ((3. / o->p.Rebs) * o->p.ub * (zu - 2.) + 0.5);
    const size_t nhalf = (size_t)xdu * ydu * zu * 19;
    int g, half;
    if (xdu < 1 || ydu < 1 || zdu < 2) {
        snprintf(err, errlen, "bad uniform grid size");
        return -1;
    }
    /* every stencil must lie inside the uniform array and its halo: x, y in [-2, n+3],
       z in [0, zdu+1]; the reference reads out of bounds otherwise */
    for (half = 0; half < 2; half++)
        for (g = 0; g < 3; g++) {
            int nx, ny, klo, khi, i, j, k;
            uni_ranges(o, g, &nx, &ny, &klo, &khi);
            for (i = 0; i < nx; i++)
                for (j = 0; j < ny; j++)
                    for (k = klo; k < khi; k++) {
                        int ind[3], jnd[3], knd[3];
                        double dx, dy, dz;
                        if (j > 0 && i > 0 && j < ny - 1 && i < nx - 1 && k > klo && k < khi - 1) continue;
                        uni_stencil(o, dxf, dxc, dxr, g, half, i, j, k, ind, jnd, knd, &dx, &dy, &dz);
                        if (ind[0] < -2 || ind[2] > xdu + 3 || jnd[0] < -2 || jnd[2] > ydu + 3 || knd[0] < 0 ||
                            knd[2] > zu - 1) {
                            snprintf(err, errlen, "uniform grid %dx%dx%d does not cover grid %d node (%d,%d,%d)",
                                     xdu, ydu, zu, g, i, j, k);
                            return -1;
                        }
                    }
        }
    /* main:986-987: init() runs before solid_init() here, so no node is marked solid
       yet and every node, the band wall layers included, starts at rest (init.c:17-52);
       the interpolation then overwrites every non-wall node.  Levels stay zero. */
    orc_init_rest(o);
    {
        int i, j, a;
        for (i = 0; i < o->Xf; i++)
            for (j = 0; j < o->Yf; j++)
                for (a = 0; a < 19; a++) {
                    o->lf[FI(o, i, j, 0) + a] = o->lft[FI(o, i, j, 0) + a] = w[a];
                    o->uf[FI(o, i, j, o->zdf - 1) + a] = o->uft[FI(o, i, j, o->zdf - 1) + a] = w[a];
                }
    }
    for (half = 0; half < 2; half++) {
        const uni_t u = {uni + half * nhalf, xdu, ydu, zu};
        double *dst[3];
        dst[0] = half ? o->cf : o->cft;
        dst[1] = half ? o->lf : o->lft;
        dst[2] = half ? o->uf : o->uft;
        for (g = 0; g < 3; g++) {
            const double fac = g == 0 ? (dxr / dxc) * (omega_u / o->tauc) : (dxr / dxf) * (omega_u / o->tauf);
            int nx, ny, klo, khi, i;
            uni_ranges(o, g, &nx, &ny, &klo, &khi);
#pragma omp parallel for schedule(static)
            for (i = 0; i < nx; i++) {
                int j, k, ind[3], jnd[3], knd[3];
                double dx, dy, dz, fi[19];
                for (j = 0; j < ny; j++)
                    for (k = klo; k < khi; k++) {
                        double *out = dst[g] + (g == 0 ? CI(o, i, j, k) : FI(o, i, j, k));
                        uni_stencil(o, dxf, dxc, dxr, g, half, i, j, k, ind, jnd, knd, &dx, &dy, &dz);
                        interp_qd(&u, ind, jnd, knd, dx, dy, dz, fi);
                        uni_rescale(fi, fac, out);
                    }
            }
        }
    }
    /* inter_init.c:150-205: the k = 1 plane fills lin, the k = zdc-2 plane's f_neq fills
       uin, but its rho and velocity are written into lin as well (the reference's
       indexing), so lin carries the upper plane's macros and uin keeps zero macros */
    {
        const size_t np = (size_t)o->Xc * o->Yc * 23;
        const size_t lim = (size_t)19 * (o->Xc + 6) * (o->Yc + 6);
        double *tl = malloc(np * 8), *tu = malloc(np * 8);
        size_t n;
        int i, j, b;
        if (!tl || !tu) {
            free(tl); free(tu);
            snprintf(err, errlen, "out of memory");
            return -1;
        }
        sample_plane(o, o->cft, 1, tl);
        sample_plane(o, o->cft, o->zdc - 2, tu);
        for (n = 0; n < np; n += 23) {
            for (b = 0; b < 4; b++) { o->lin[n + b] = tu[n + b]; o->uin[n + b] = 0.; }
            for (b = 4; b < 23; b++) { o->lin[n + b] = tl[n + b]; o->uin[n + b] = tu[n + b]; }
        }
        /* inter_init.c:245-257: the copy into the older levels runs over the ghosted
           (Xc+6) x (Yc+6) planes with a stride of 19 where they have 23, so it copies
           the first 19/23 of each plane's memory and leaves the rest zero */
        for (i = 0; i < o->Xc; i++)
            for (j = 0; j < o->Yc; j++)
                for (b = 0; b < 23; b++) {
                    const size_t at = PI(o, i, j) + b;
                    const size_t flat = ((size_t)(i + 2) * (o->Yc + 6) + (size_t)(j + 2)) * 23 + b;
                    const int copied = flat < lim;
                    o->linp[at] = o->linpp[at] = copied ? o->lin[at] : 0.;
                    o->uinp[at] = o->uinpp[at] = copied ? o->uin[at] : 0.;
                }
        free(tl);
        free(tu);
    }
    return 0;
}

/* a synthetic uniform-grid field for the '-i' start, [2][xdu][ydu][zdu+2][19]: each
 * half is the equilibrium of a laminar profile with seeded noise in rho and u
 * (different noise in the two halves), wall layers k = 0, zdu+1 at rest */
void orc_uniform_field(int xdu, int ydu, int zdu, uint64_t seed, double u_max, double noise, double *out) {
    const int zu = zdu + 2;
    const size_t nhalf = (size_t)xdu * ydu * zu * 19;
    int half;
    for (half = 0; half < 2; half++) {
        size_t node;
        for (node = 0; node < nhalf / 19; node++) {
            const int k = (int)(node % zu);
            double *f = out + half * nhalf + node * 19;
            double dens = 1., ueq[3] = {0., 0., 0.}, v;
            int a;
            if (k > 0 && k < zu - 1) {
                const double zeta = (k - 0.5) / (double)(zu - 2), amp = noise * u_max;
                dens = 1. + 0.01 * noise * orc_prng_sym(seed, 3 + half, 3, node);
                ueq[0] = 4. * u_max * zeta * (1. - zeta) + amp * orc_prng_sym(seed, 3 + half, 0, node);
                ueq[1] = amp * orc_prng_sym(seed, 3 + half, 1, node);
                ueq[2] = amp * orc_prng_sym(seed, 3 + half, 2, node);
            }
            v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
            for (a = 0; a < 19; a++) f[a] = feq_of(a, dens, ueq, v);
        }
    }
}

/* ---- interface sampling: xcomputations.c:558-625 ------------------------------ */
static void sample_plane(const orc_t *o, const double *f, int k, double *plane) {
    int i, j, a;
#pragma omp parallel for private(j, a) schedule(static)
    for (i = 0; i < o->Xc; i++)
        for (j = 0; j < o->Yc; j++) {
            const double *fp = f + CI(o, i, j, k);
            double dens, ueq[3], v;
            double *pl = plane + PI(o, i, j);
            moments(fp, &dens, ueq);
            v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
            pl[0] = dens;
            pl[1] = ueq[0];
            pl[2] = ueq[1];
            pl[3] = ueq[2];
            for (a = 0; a < 19; a++) {
                double feq = feq_of(a, dens, ueq, v);
                pl[a + 4] = fp[a] - feq;
            }
        }
}

/* ---- coarse collide + stream + sampling: xcomputations.c:16-950 --------------- */
void orc_xcomputations(orc_t *o) {
    const double Gx = o->force[0]; /* :48 */
    const double tauc = o->tauc;
    int i;
    /* collide every coarse node in place (all codes are 0 or 1: :49-134, :291-342) */
#pragma omp parallel for schedule(static)
    for (i = 0; i < o->Xc; i++) {
        int j, k;
        for (j = 0; j < o->Yc; j++)
            for (k = 0; k < o->zdc; k++) collide(o->cft + CI(o, i, j, k), Gx, tauc);
    }
    /* push-stream ftemp -> f with periodic x/y (halo :206-546) and bounded z at
       k=0, zdc-1 (code 1: :153-165, :372-384) */
#pragma omp parallel for schedule(static)
    for (i = 0; i < o->Xc; i++) {
        int j, k, a;
        for (j = 0; j < o->Yc; j++)
            for (k = 0; k < o->zdc; k++) {
                const double *src = o->cft + CI(o, i, j, k);
                for (a = 0; a < 19; a++) {
                    int kto = k + e2[a];
                    if (kto < 0 || kto >= o->zdc) continue;
                    o->cf[CI(o, wrap(i + e0[a], o->Xc), wrap(j + e1[a], o->Yc), kto) + a] = src[a];
                }
            }
    }
    /* rotate the interface time levels: :548-556 */
    {
        double *t = o->linpp; o->linpp = o->linp; o->linp = o->lin; o->lin = t;
        t = o->uinpp; o->uinpp = o->uinp; o->uinp = o->uin; o->uin = t;
    }
    /* sample k=1 -> lin, k=zdc-2 -> uin: :558-625 */
    sample_plane(o, o->cf, 1, o->lin);
    sample_plane(o, o->cf, o->zdc - 2, o->uin);
}

/* ---- fine collide + stream with walls: xfcomputations.c:16-1144 --------------- */
static void band_push(orc_t *o, const double *ft, double *f, const int8_t *wall, int upper) {
    const int wallk = upper ? o->zdf - 1 : 0;
    const int k0 = upper ? 0 : 1, k1 = upper ? o->zdf - 1 : o->zdf;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < o->Xf; i++) {
        int j, k, a;
        for (j = 0; j < o->Yf; j++)
            for (k = k0; k < k1; k++) {
                const double *src = ft + FI(o, i, j, k);
                for (a = 0; a < 19; a++) {
                    int kto = k + e2[a];
                    int ito, jto, sld, sld1;
                    if (kto < 0 || kto >= o->zdf) continue; /* bounded layer: :252, :761 */
                    ito = wrap(i + e0[a], o->Xf);
                    jto = wrap(j + e1[a], o->Yf);
                    if (kto != wallk) { /* sld > -1: :232-233 */
                        f[FI(o, ito, jto, kto) + a] = src[a];
                        continue;
                    }
                    sld = wall[(size_t)ito * o->Yf + jto];  /* :230 */
                    sld1 = wall[(size_t)i * o->Yf + j];     /* :231 */
                    if ((sld == -1) && (sld1 == -1))        /* halfway bounce-back :234-235 */
                        f[FI(o, i, j, k) + opposite[a]] = src[a];
                    else if ((sld == -1) && (sld1 == -2))   /* specular :236-237 */
                        f[FI(o, ito, jto, k) + mirror[a]] = src[a];
                    else if (sld == -2)                      /* specular :238-239 */
                        f[FI(o, ito, jto, k) + mirror[a]] = src[a];
                }
            }
    }
}

void orc_xfcomputations(orc_t *o) {
    const double Gx = (1. / ((double)o->GR)) * o->force[0]; /* :42 with GR = 1./GR (main:1349) */
    const double tauf = o->tauf;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < o->Xf; i++) {
        int j, k;
        for (j = 0; j < o->Yf; j++)
            for (k = 0; k < o->zdf; k++) {
                if (k != o->zdf - 1) collide(o->uft + FI(o, i, j, k), Gx, tauf); /* s>-1: :43-128 */
                if (k != 0) collide(o->lft + FI(o, i, j, k), Gx, tauf);         /* :129-214 */
            }
    }
    band_push(o, o->uft, o->uf, o->wu, 1);
    band_push(o, o->lft, o->lf, o->wl, 0);
}

/* ---- Hermite in time: tinterpolate_hrmt.c:16-55 ------------------------------- */
static void hermite(const orc_t *o, const double *ln, const double *lp, const double *lpp,
                    double *out, double tmp) {
    size_t n, np = (size_t)o->Xc * o->Yc * 23;
    for (n = 0; n < np; n++) {
        double c1 = lp[n];
        double c2 = (ln[n] - lpp[n]) * 0.5;
        double c3 = ln[n];
        double c4 = (3. * ln[n] - 4. * lp[n] + lpp[n]) * 0.5;
        out[n] = ((2. * tmp * tmp * tmp - 3. * tmp * tmp + 1.) * c1 +
                  (tmp * tmp * tmp - 2. * tmp * tmp + tmp) * c2 +
                  (-2. * tmp * tmp * tmp + 3. * tmp * tmp) * c3 + (tmp * tmp * tmp - tmp * tmp) * c4);
    }
}

void orc_tinterpolate_hrmt(orc_t *o, int iter) {
    double tmp = (((double)iter) / ((double)o->GR)); /* :20 */
    hermite(o, o->lin, o->linp, o->linpp, o->lint, tmp);
    hermite(o, o->uin, o->uinp, o->uinpp, o->uint_, tmp);
}

/* ---- 4th-order derivatives: derivative.c:17-89 (periodic in x and y) ---------- */
void orc_derivative(orc_t *o, double *addr) {
    const int X = o->Xc, Y = o->Yc;
    double *s0 = addr, *s1 = addr + SLOT(o, 1), *s2 = addr + SLOT(o, 2), *s3 = addr + SLOT(o, 3);
    int i;
#define V(s, i, j, a) (s)[PI(o, wrap((i), X), wrap((j), Y)) + (a)]
#pragma omp parallel for schedule(static)
    for (i = 0; i < X; i++) {
        int j, a;
        for (j = 0; j < Y; j++)
            for (a = 0; a < 23; a++) {
                V(s1, i, j, a) = (2. / 3.) * (V(s0, i + 1, j, a) - V(s0, i - 1, j, a)) -
                                 (1. / 12.) * (V(s0, i + 2, j, a) - V(s0, i - 2, j, a)); /* :118 */
                V(s2, i, j, a) = (2. / 3.) * (V(s0, i, j + 1, a) - V(s0, i, j - 1, a)) -
                                 (1. / 12.) * (V(s0, i, j + 2, a) - V(s0, i, j - 2, a)); /* :119 */
            }
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < X; i++) {
        int j, a;
        for (j = 0; j < Y; j++)
            for (a = 0; a < 23; a++)
                V(s3, i, j, a) = (2. / 3.) * (V(s2, i + 1, j, a) - V(s2, i - 1, j, a)) -
                                 (1. / 12.) * (V(s2, i + 2, j, a) - V(s2, i - 2, j, a)); /* :139 */
    }
#undef V
}

/* ---- bicubic coefficients: bcucof_opt.c:16-62 --------------------------------- */
static const int wt_d[16][16] = {
    {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
    {-3, 0, 0, 3, 0, 0, 0, 0, -2, 0, 0, -1, 0, 0, 0, 0},
    {2, 0, 0, -2, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0},
    {0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0},
    {0, 0, 0, 0, -3, 0, 0, 3, 0, 0, 0, 0, -2, 0, 0, -1},
    {0, 0, 0, 0, 2, 0, 0, -2, 0, 0, 0, 0, 1, 0, 0, 1},
    {-3, 3, 0, 0, -2, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, -3, 3, 0, 0, -2, -1, 0, 0},
    {9, -9, 9, -9, 6, 3, -3, -6, 6, -6, -3, 3, 4, 2, 1, 2},
    {-6, 6, -6, 6, -4, -2, 2, 4, -3, 3, 3, -3, -2, -1, -1, -2},
    {2, -2, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 2, -2, 0, 0, 1, 1, 0, 0},
    {-6, 6, -6, 6, -3, -3, 3, 3, -4, 4, 2, -2, -2, -2, -1, -1},
    {4, -4, 4, -4, 2, 2, -2, -2, 2, -2, -2, 2, 1, 1, 1, 1}};

void orc_bcucof(const orc_t *o, const double *addr, int ic, int jc, double *c) {
    const double dx = 1.; /* :44 */
    int in = wrap(ic + 1, o->Xc), jn = wrap(jc + 1, o->Yc);
    size_t s1 = SLOT(o, 1), s2 = SLOT(o, 2), s3 = SLOT(o, 3);
    size_t p00 = PI(o, ic, jc), p10 = PI(o, in, jc), p11 = PI(o, in, jn), p01 = PI(o, ic, jn);
    int a, i, k;
    for (a = 0; a < 23; a++) { /* :47-53 */
        double x[16], xx;
        x[0] = addr[p00 + a]; x[4] = dx * addr[s1 + p00 + a]; x[8] = dx * addr[s2 + p00 + a]; x[12] = dx * dx * addr[s3 + p00 + a];
        x[1] = addr[p10 + a]; x[5] = dx * addr[s1 + p10 + a]; x[9] = dx * addr[s2 + p10 + a]; x[13] = dx * dx * addr[s3 + p10 + a];
        x[2] = addr[p11 + a]; x[6] = dx * addr[s1 + p11 + a]; x[10] = dx * addr[s2 + p11 + a]; x[14] = dx * dx * addr[s3 + p11 + a];
        x[3] = addr[p01 + a]; x[7] = dx * addr[s1 + p01 + a]; x[11] = dx * addr[s2 + p01 + a]; x[15] = dx * dx * addr[s3 + p01 + a];
        for (i = 0; i <= 15; i++) { /* :55-61 */
            xx = 0.0;
            for (k = 0; k <= 15; k++) xx += wt_d[i][k] * x[k];
            c[a * 16 + i] = xx;
        }
    }
}

/* ---- coarse -> fine interface: ctfdtransfer_opt.c:17-175 ----------------------- */
static void ctfd_band(orc_t *o, const double *addr, double *f, double *ft, int kf) {
    const int GR = o->GR;
    const double rescale_num = o->tauc, rescale_den = (double)GR * o->tauf;
    int ic;
#pragma omp parallel for schedule(static)
    for (ic = 0; ic < o->Xc; ic++) {
        double c[23 * 16];
        int jc, ii, jj, nn, a;
        for (jc = 0; jc < o->Yc; jc++) {
            orc_bcucof(o, addr, ic, jc, c); /* :35 */
            for (ii = 0; ii < GR; ii++)
                for (jj = 0; jj < GR; jj++) {
                    int i = ic * GR + ii, j = jc * GR + jj;
                    double tmp, tt, uu, dens, ueq[3], v;
                    double *fp = f + FI(o, i, j, kf), *ftp = ft + FI(o, i, j, kf);
                    tmp = i % GR; tt = ((double)tmp) / ((double)GR); /* :55-56 */
                    tmp = j % GR; uu = ((double)tmp) / ((double)GR);
#define CC(q, n, m) c[(q) * 16 + (n) * 4 + (m)]
                    dens = 0.;
                    for (nn = 3; nn > -1; nn--)
                        dens = tt * dens + ((CC(0, nn, 3) * uu + CC(0, nn, 2)) * uu + CC(0, nn, 1)) * uu + CC(0, nn, 0);
                    ueq[0] = 0.;
                    for (nn = 3; nn > -1; nn--)
                        ueq[0] = tt * ueq[0] + ((CC(1, nn, 3) * uu + CC(1, nn, 2)) * uu + CC(1, nn, 1)) * uu + CC(1, nn, 0);
                    ueq[1] = 0.;
                    for (nn = 3; nn > -1; nn--)
                        ueq[1] = tt * ueq[1] + ((CC(2, nn, 3) * uu + CC(2, nn, 2)) * uu + CC(2, nn, 1)) * uu + CC(2, nn, 0);
                    ueq[2] = 0.;
                    for (nn = 3; nn > -1; nn--)
                        ueq[2] = tt * ueq[2] + ((CC(3, nn, 3) * uu + CC(3, nn, 2)) * uu + CC(3, nn, 1)) * uu + CC(3, nn, 0);
                    v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
                    for (a = 0; a < 19; a++) {
                        double feq = feq_of(a, dens, ueq, v), fneq = 0.;
                        for (nn = 3; nn > -1; nn--)
                            fneq = tt * fneq + ((CC(a + 4, nn, 3) * uu + CC(a + 4, nn, 2)) * uu + CC(a + 4, nn, 1)) * uu + CC(a + 4, nn, 0);
                        fp[a] = feq + (rescale_num / rescale_den) * fneq;  /* :93 */
                        ftp[a] = feq + (rescale_num / rescale_den) * fneq; /* :94 */
                    }
#undef CC
                }
        }
    }
}

void orc_ctfdtransfer(orc_t *o, int iter) {
    (void)iter;
    ctfd_band(o, o->lint, o->lf, o->lft, o->zdf - 1); /* lower band k = zdf-1: :31-99 */
    ctfd_band(o, o->uint_, o->uf, o->uft, 0);         /* upper band k = 0: :102-171 */
}

/* ---- fine -> coarse injection: ftcdtransfer.c:16-118 -------------------------- */
static void ftcd_side(orc_t *o, const double *fsrc, int kf, int kc) {
    const double fac = o->GR * (o->tauf / o->tauc); /* GR*(tauf/tauc): :66 */
    int ic;
#pragma omp parallel for schedule(static)
    for (ic = 0; ic < o->Xc; ic++) {
        int jc, a;
        for (jc = 0; jc < o->Yc; jc++) {
            const double *fp = fsrc + FI(o, ic * o->GR, jc * o->GR, kf);
            double dens, ueq[3], v;
            double *cf = o->cf + CI(o, ic, jc, kc), *cft = o->cft + CI(o, ic, jc, kc);
            moments(fp, &dens, ueq);
            v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
            for (a = 0; a < 19; a++) {
                double feq = feq_of(a, dens, ueq, v);
                double fneq = fp[a] - feq;
                cf[a] = feq + fac * fneq;
                cft[a] = feq + fac * fneq;
            }
        }
    }
}

/* ---- filtered injection: ftcdtransfer_lesfltr.c:16-381 (FLTR) --------------------------
 * f_neq of the 5 fine layers kf-2 .. kf+2 around the coincident layer (:160-200), box-
 * filtered over 5 x 5 x 5 nodes with the trapezoid weights of :22-28 and summed term by
 * term as (w f_neq) / GR^3 in ii, jj, kk order (:222-236); rho and the velocity of the
 * equilibrium come unfiltered from the coincident node, where the reference stores the
 * raw first moments sum e f (not divided by rho) and uses them as the velocity
 * (:173-176, :207-212).  Only the post-stream coarse array is written (:242). */
static void ftcd_flt_side(orc_t *o, const double *fsrc, int kf, int kc) {
    const int GR = o->GR, Xf = o->Xf, Yf = o->Yf;
    const double fac = GR * (o->tauf / o->tauc);
    const double den = (double)GR * GR * GR;
    static const double wl[5] = {0.5, 1., 1., 1., 0.5}; /* wght[ii][jj][kk] = wl[ii] wl[jj] wl[kk] */
    double *slab = malloc(sizeof(double) * (size_t)Xf * Yf * 5 * 23);
    int i, ic;
#pragma omp parallel for schedule(static)
    for (i = 0; i < Xf; i++) {
        int j, kk, a;
        for (j = 0; j < Yf; j++)
            for (kk = -2; kk < 3; kk++) {
                const double *fp = fsrc + FI(o, i, j, kf + kk);
                double *s = slab + (((size_t)i * Yf + j) * 5 + (kk + 2)) * 23;
                double ux = 0., uy = 0., uz = 0., dens = 0., ueq[3], v;
                for (a = 0; a < 19; a++) {
                    ux += E0[a] * fp[a];
                    uy += E1[a] * fp[a];
                    uz += E2[a] * fp[a];
                    dens += fp[a];
                }
                ueq[0] = ux / dens;
                ueq[1] = uy / dens;
                ueq[2] = uz / dens;
                v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
                s[0] = dens;
                s[1] = ux;
                s[2] = uy;
                s[3] = uz;
                for (a = 0; a < 19; a++) s[4 + a] = fp[a] - feq_of(a, dens, ueq, v);
            }
    }
#pragma omp parallel for schedule(static)
    for (ic = 0; ic < o->Xc; ic++) {
        int jc, a, ii, jj, kk;
        for (jc = 0; jc < o->Yc; jc++) {
            const int i = ic * GR, j = jc * GR;
            const double *c = slab + (((size_t)i * Yf + j) * 5 + 2) * 23;
            const double dens = c[0], ueq[3] = {c[1], c[2], c[3]};
            const double v = ueq[0] * ueq[0] + ueq[1] * ueq[1] + ueq[2] * ueq[2];
            double fneqt[19], *cf = o->cf + CI(o, ic, jc, kc);
            for (a = 0; a < 19; a++) fneqt[a] = 0.;
            for (ii = -2; ii < 3; ii++)
                for (jj = -2; jj < 3; jj++)
                    for (kk = -2; kk < 3; kk++) {
                        const int in = (i + ii + Xf) % Xf, jn = (j + jj + Yf) % Yf;
                        const double *s = slab + (((size_t)in * Yf + jn) * 5 + (kk + 2)) * 23;
                        const double wt = wl[ii + 2] * wl[jj + 2] * wl[kk + 2];
                        for (a = 0; a < 19; a++) fneqt[a] += (wt * s[4 + a]) / den;
                    }
            for (a = 0; a < 19; a++) cf[a] = feq_of(a, dens, ueq, v) + fac * fneqt[a];
        }
    }
    free(slab);
}

void orc_ftcdtransfer(orc_t *o) {
    if (o->p.fltr) {  /* main:1369-1371 */
        ftcd_flt_side(o, o->lf, o->zdf - 1 - o->GR, 0);
        ftcd_flt_side(o, o->uf, o->GR, o->zdc - 1);
        return;
    }
    ftcd_side(o, o->lf, o->zdf - 1 - o->GR, 0);     /* :34, :66-67 */
    ftcd_side(o, o->uf, o->GR, o->zdc - 1);         /* :79, :111-112 */
}

void orc_exchange(orc_t *o) { double *t = o->cf; o->cf = o->cft; o->cft = t; } /* exchange.c:16-25 */
void orc_fexchange(orc_t *o) {                                               /* fexchange.c:15-28 */
    double *t = o->uf; o->uf = o->uft; o->uft = t;
    t = o->lf; o->lf = o->lft; o->lft = t;
}

/* ---- flux control and logs: encalc2.c:17-147 ------------------------------------ */
void orc_encalc2(orc_t *o, long ss, dnsg_diag *dg) {
    const int GR = o->GR, zdf = o->zdf, zdc = o->zdc, NZs = o->NZs;
    double momentum = 0., denst = 0., cmomentum = 0., cdenst = 0., coef, dens, u[3];
    int i, j, k;
    for (i = 0; i < o->Xf; i++)               /* :33-56 */
        for (j = 0; j < o->Yf; j++)
            for (k = 1; k < (zdf - GR); k++) {
                const double *fp = o->lft + FI(o, i, j, k);
                double ux = 0.;
                int a;
                dens = 0.;
                for (a = 0; a < 19; a++) { dens += fp[a]; ux += E0[a] * fp[a]; }
                coef = (k == (zdf - GR)) ? 0.5 : 1.;
                momentum += coef * ux;
                denst += coef * dens;
            }
    for (i = 0; i < o->Xc; i++)               /* :58-83 */
        for (j = 0; j < o->Yc; j++)
            for (k = 0; k < zdc; k++) {
                const double *fp = o->cft + CI(o, i, j, k);
                double ux = 0.;
                int a;
                dens = 0.;
                for (a = 0; a < 19; a++) { dens += fp[a]; ux += E0[a] * fp[a]; }
                if ((k == 0) || (k == (zdc - 1))) coef = 0.5 * GR * GR * GR;
                else coef = 1. * GR * GR * GR;
                momentum += coef * ux;
                cmomentum += coef * ux;
                denst += coef * dens;
                cdenst += coef * dens;
            }
    for (i = 0; i < o->Xf; i++)               /* :85-107 */
        for (j = 0; j < o->Yf; j++)
            for (k = GR; k < (zdf - 1); k++) {
                const double *fp = o->uft + FI(o, i, j, k);
                double ux = 0.;
                int a;
                dens = 0.;
                for (a = 0; a < 19; a++) { dens += fp[a]; ux += E0[a] * fp[a]; }
                coef = (k == GR) ? 0.5 : 1.;
                momentum += coef * ux;
                denst += coef * dens;
            }
    {
        double vol = ((double)((o->Xf + 1 - 1.) * (o->Yf + 1 - 1.) * 1 * 1 * (NZs - 2.))); /* :109 */
        momentum /= vol;
        cmomentum /= vol;
        denst /= vol;
        cdenst /= vol;
    }
    (void)u;
    o->rho[0] = denst; /* :125-126 */
    o->rho[1] = cdenst;
    if (o->p.force_mode == DNSG_FORCE_FLUX) o->force[0] = o->p.ub - momentum; /* :129 */
    else o->force[0] = o->p.g_const;
    o->force[1] = cmomentum;
    if (dg) {
        dg->time = ((double)o->p.TSTR + o->d.dt * ss);
        dg->force = o->force[0];
        dg->momentum = momentum;
        dg->density = denst;
        dg->flowrate = momentum * ((o->Yf) * (NZs - 2)) / denst; /* :138 */
    }
}

/* ---- wall shear stress: wshearstr.c:15-168 -------------------------------------- */
int orc_wshearstr(orc_t *o, long ss, dnsg_diag *dg) {
    const int zdf = o->zdf, NZs = o->NZs, GR = o->GR;
    double mu = (1. / 6.) * (2. / o->tauf - 1); /* :35 */
    double stdown = 0., stup = 0., davg = 0., pgrad, stress;
    const double area = ((double)((o->Xf + 1 - 1.) * (o->Yf + 1 - 1.) * 1 * 1));
    int i, j;
#define UX_AT(arr, kk, out, densout)                              \
    do {                                                          \
        const double *fp_ = (arr) + FI(o, i, j, (kk));            \
        double d_ = 0., x_ = 0.;                                  \
        int a_;                                                   \
        for (a_ = 0; a_ < 19; a_++) { d_ += fp_[a_]; x_ += E0[a_] * fp_[a_]; } \
        (out) = x_ / d_;                                          \
        (densout) = d_;                                           \
    } while (0)
    for (i = 0; i < o->Xf; i++)            /* :37-70 */
        for (j = 0; j < o->Yf; j++) {
            double ux1, ux2, ux3, dens;
            UX_AT(o->lft, 1, ux1, dens);
            davg += dens;
            UX_AT(o->lft, 2, ux2, dens);
            UX_AT(o->lft, 3, ux3, dens);
            stdown += (-2. * ux1 + 3. * ux2 - ux3);
        }
    davg /= area;   /* :72 */
    stdown /= area; /* :73 */
    stdown *= mu * davg; /* :81 */
    davg = 0.;
    for (i = 0; i < o->Xf; i++)            /* :85-118 */
        for (j = 0; j < o->Yf; j++) {
            double ux1, ux2, ux3, dens;
            UX_AT(o->uft, zdf - 2, ux1, dens);
            UX_AT(o->uft, zdf - 3, ux2, dens);
            UX_AT(o->uft, zdf - 4, ux3, dens);
            stup += -(2. * ux1 - 3. * ux2 + ux3);
            davg += dens;                  /* density of layer zdf-4 (quirk) */
        }
#undef UX_AT
    davg /= area;
    stup /= area;
    pgrad = o->force[0];                    /* :130-132 */
    pgrad = pgrad * (NZs - 2.) / (2. * GR);
    stup *= mu * davg;                      /* :134 */
    stress = 0.5 * (stup + stdown);         /* :138 */
    if (dg) {
        dg->time = ((double)o->p.TSTR + o->d.dt * ss);
        dg->str_up = stup;
        dg->str_down = stdown;
        dg->stress = stress;
        dg->pgrad = pgrad;
    }
    return isnan(stress) ? DNSG_ERR_NAN : 0;
}

/* ---- mass check: main:1193-1250 --------------------------------------------------- */
double orc_mass(const orc_t *o) {
    const int GR = o->GR, zdf = o->zdf, zdc = o->zdc;
    double rho = 0., rz, coef;
    int i, j, k, a;
    for (i = 0; i < o->Xf; i++)
        for (j = 0; j < o->Yf; j++)
            for (k = 0; k < zdf; k++) {
                rz = 0.;
                for (a = 0; a < 19; a++) rz += o->lft[FI(o, i, j, k) + a];
                coef = (k == (zdf - 1)) ? 0.5 : 1.;
                rho += (coef * rz);
            }
    for (i = 0; i < o->Xc; i++)
        for (j = 0; j < o->Yc; j++)
            for (k = 1; k < (zdc - 1); k++) {
                rz = 0.;
                for (a = 0; a < 19; a++) rz += o->cft[CI(o, i, j, k) + a];
                if ((k == 1) || (k == (zdc - 2))) coef = 0.5 * GR * GR * GR;
                else coef = 1. * GR * GR * GR;
                rho += (coef * rz);
            }
    for (i = 0; i < o->Xf; i++)
        for (j = 0; j < o->Yf; j++)
            for (k = 0; k < zdf; k++) {
                rz = 0.;
                for (a = 0; a < 19; a++) rz += o->uft[FI(o, i, j, k) + a];
                coef = (k == 0) ? 0.5 : 1.;
                rho += (coef * rz);
            }
    rho /= ((double)((o->Xf + 1 - 1.) * (o->Yf + 1 - 1.) * 1 * 1 * (o->NZs - 2.))); /* :1247 */
    return rho;
}

/* ---- plane statistics: statistics.c:15-300 --------------------------------------- */
void orc_statistics(const orc_t *o, double *out) {
    const int zdf = o->zdf, zdc = o->zdc, NZ = o->NZ;
    double *UAVG = out, *VAVG = out + NZ, *WAVG = out + 2 * NZ, *UTIN = out + 3 * NZ,
           *VTIN = out + 4 * NZ, *WTIN = out + 5 * NZ, *REYS = out + 6 * NZ, *DAVG = out + 7 * NZ,
           *DTIN = out + 8 * NZ;
    int i, j, k, a, extx, exty;
    memset(out, 0, sizeof(double) * 9 * NZ);
#define NODE_UVW(fp)                                                                  \
    u = 0.; v = 0.; ww = 0.; dens = 0.;                                                \
    for (a = 0; a < 19; a++) { dens += (fp)[a]; u += E0[a] * (fp)[a]; v += E1[a] * (fp)[a]; ww += E2[a] * (fp)[a]; }
    {
        double u, v, ww, dens;
        extx = o->Xf; exty = o->Yf;
        for (i = 0; i < extx; i++)                      /* :48-69 */
            for (j = 0; j < exty; j++)
                for (k = 1; k < zdf; k++) {
                    NODE_UVW(o->lft + FI(o, i, j, k));
                    UAVG[k] += (u / dens); VAVG[k] += (v / dens); WAVG[k] += (ww / dens); DAVG[k] += dens;
                }
        for (k = 0; k < zdf; k++) {                     /* :71-77 */
            UAVG[k] /= (double)(extx * exty * 1 * 1); VAVG[k] /= (double)(extx * exty * 1 * 1);
            WAVG[k] /= (double)(extx * exty * 1 * 1); DAVG[k] /= (double)(extx * exty * 1 * 1);
        }
        extx = o->Xc; exty = o->Yc;
        for (i = 0; i < extx; i++)                      /* :82-103 */
            for (j = 0; j < exty; j++)
                for (k = 2; k < (zdc - 2); k++) {
                    NODE_UVW(o->cft + CI(o, i, j, k));
                    UAVG[zdf - 2 + k] += (u / dens); VAVG[zdf - 2 + k] += (v / dens);
                    WAVG[zdf - 2 + k] += (ww / dens); DAVG[zdf - 2 + k] += dens;
                }
        for (k = 2; k < (zdc - 2); k++) {               /* :105-111 */
            UAVG[zdf - 2 + k] /= (double)(extx * exty * 1 * 1); VAVG[zdf - 2 + k] /= (double)(extx * exty * 1 * 1);
            WAVG[zdf - 2 + k] /= (double)(extx * exty * 1 * 1); DAVG[zdf - 2 + k] /= (double)(extx * exty * 1 * 1);
        }
        extx = o->Xf; exty = o->Yf;
        for (i = 0; i < extx; i++)                      /* :116-137 */
            for (j = 0; j < exty; j++)
                for (k = 0; k < (zdf - 1); k++) {
                    NODE_UVW(o->uft + FI(o, i, j, k));
                    UAVG[zdf + zdc - 4 + k] += (u / dens); VAVG[zdf + zdc - 4 + k] += (v / dens);
                    WAVG[zdf + zdc - 4 + k] += (ww / dens); DAVG[zdf + zdc - 4 + k] += dens;
                }
        for (k = 0; k < zdf; k++) {                     /* :139-145 */
            UAVG[zdf + zdc - 4 + k] /= (double)(extx * exty * 1 * 1); VAVG[zdf + zdc - 4 + k] /= (double)(extx * exty * 1 * 1);
            WAVG[zdf + zdc - 4 + k] /= (double)(extx * exty * 1 * 1); DAVG[zdf + zdc - 4 + k] /= (double)(extx * exty * 1 * 1);
        }
    }
    {
        double u, v, ww, dens, UF, VF, WF, DF;
        extx = o->Xf; exty = o->Yf;
        for (i = 0; i < extx; i++)                      /* :162-188 */
            for (j = 0; j < exty; j++)
                for (k = 1; k < zdf; k++) {
                    NODE_UVW(o->lft + FI(o, i, j, k));
                    UF = (u / dens) - UAVG[k]; VF = (v / dens) - VAVG[k]; WF = (ww / dens) - WAVG[k]; DF = dens - DAVG[k];
                    UTIN[k] += UF * UF; VTIN[k] += VF * VF; WTIN[k] += WF * WF; REYS[k] += UF * WF; DTIN[k] += DF * DF;
                }
        for (k = 0; k < zdf; k++) {                     /* :190-197 */
            UTIN[k] /= (double)(extx * exty * 1 * 1); VTIN[k] /= (double)(exty * extx * 1 * 1);
            WTIN[k] /= (double)(exty * extx * 1 * 1); REYS[k] /= (double)(extx * exty * 1 * 1);
            DTIN[k] /= (double)(extx * exty * 1 * 1);
        }
        extx = o->Xc; exty = o->Yc;
        for (i = 0; i < extx; i++)                      /* :202-228 */
            for (j = 0; j < exty; j++)
                for (k = 2; k < (zdc - 2); k++) {
                    int s = zdf - 2 + k;
                    NODE_UVW(o->cft + CI(o, i, j, k));
                    UF = (u / dens) - UAVG[s]; VF = (v / dens) - VAVG[s]; WF = (ww / dens) - WAVG[s]; DF = dens - DAVG[s];
                    UTIN[s] += UF * UF; VTIN[s] += VF * VF; WTIN[s] += WF * WF; REYS[s] += UF * WF; DTIN[s] += DF * DF;
                }
        for (k = 2; k < (zdc - 2); k++) {               /* :230-237 */
            int s = zdf - 2 + k;
            UTIN[s] /= (double)(extx * exty * 1 * 1); VTIN[s] /= (double)(exty * extx * 1 * 1);
            WTIN[s] /= (double)(exty * extx * 1 * 1); REYS[s] /= (double)(extx * exty * 1 * 1);
            DTIN[s] /= (double)(extx * exty * 1 * 1);
        }
        extx = o->Xf; exty = o->Yf;
        for (i = 0; i < extx; i++)                      /* :242-268 */
            for (j = 0; j < exty; j++)
                for (k = 0; k < (zdf - 1); k++) {
                    int s = zdf + zdc - 4 + k;
                    NODE_UVW(o->uft + FI(o, i, j, k));
                    UF = (u / dens) - UAVG[s]; VF = (v / dens) - VAVG[s]; WF = (ww / dens) - WAVG[s]; DF = dens - DAVG[s];
                    UTIN[s] += UF * UF; VTIN[s] += VF * VF; WTIN[s] += WF * WF; REYS[s] += UF * WF; DTIN[s] += DF * DF;
                }
        for (k = 0; k < zdf; k++) {                     /* :270-277 */
            int s = zdf + zdc - 4 + k;
            UTIN[s] /= (double)(extx * exty * 1 * 1); VTIN[s] /= (double)(exty * extx * 1 * 1);
            WTIN[s] /= (double)(exty * extx * 1 * 1); REYS[s] /= (double)(extx * exty * 1 * 1);
            DTIN[s] /= (double)(extx * exty * 1 * 1);
        }
    }
#undef NODE_UVW
    for (k = 0; k < NZ; k++) {                          /* :294-300 */
        UTIN[k] = sqrt(UTIN[k]); VTIN[k] = sqrt(VTIN[k]); WTIN[k] = sqrt(WTIN[k]); DTIN[k] = sqrt(DTIN[k]);
    }
}


/* ---- velocity/density snapshots: vel_print.c:15-274 ---------------------------------- */
static int vel_grid(const double *f, int X, int Y, int Z, const char *path_v, const char *path_d) {
    size_t n = (size_t)X * Y * Z, m;
    double *vel = malloc(n * 3 * 8), *den = malloc(n * 8);
    FILE *fv, *fd;
    if (!vel || !den) { free(vel); free(den); return DNSG_ERR_IO; }
    for (m = 0; m < n; m++) {
        const double *fp = f + m * 19;
        double ux = 0., uy = 0., uz = 0., dens = 0.;
        int a;
        for (a = 0; a < 19; a++) {
            ux += E0[a] * fp[a]; uy += E1[a] * fp[a]; uz += E2[a] * fp[a]; dens += fp[a];
        }
        if (dens > 0.) { ux /= dens; uy /= dens; uz /= dens; }   /* :99-110 */
        else { ux = 0.; uy = 0.; uz = 0.; dens = 0.; }
        vel[3 * m] = ux; vel[3 * m + 1] = uy; vel[3 * m + 2] = uz;
        den[m] = dens;
    }
    fv = fopen(path_v, "wb"); fd = fopen(path_d, "wb");
    if (fv) { fwrite(vel, 8, n * 3, fv); fclose(fv); }
    if (fd) { fwrite(den, 8, n, fd); fclose(fd); }
    free(vel); free(den);
    return (fv && fd) ? 0 : DNSG_ERR_IO;
}

int orc_vel_print(const orc_t *o, const char *dir, long cntr) {
    char pv[4096], pd[4096];
    int rc = 0;
    snprintf(pv, sizeof pv, "%s/vel-c.%.4d", dir, (int)cntr);   snprintf(pd, sizeof pd, "%s/den-c.%.4d", dir, (int)cntr);
    rc |= vel_grid(o->cft, o->Xc, o->Yc, o->zdc, pv, pd);
    snprintf(pv, sizeof pv, "%s/vel-suf.%.4d", dir, (int)cntr); snprintf(pd, sizeof pd, "%s/den-suf.%.4d", dir, (int)cntr);
    rc |= vel_grid(o->uft, o->Xf, o->Yf, o->zdf, pv, pd);
    snprintf(pv, sizeof pv, "%s/vel-slf.%.4d", dir, (int)cntr); snprintf(pd, sizeof pd, "%s/den-slf.%.4d", dir, (int)cntr);
    rc |= vel_grid(o->lft, o->Xf, o->Yf, o->zdf, pv, pd);
    return rc;
}

/* ---- mean streamwise profile without overlap: output_twoD2.c:15-169 ------------------ */
int orc_profile_len(const orc_t *o) { return 2 * (o->zdf - o->GR - 1) + o->zdc; }

void orc_profile2d(const orc_t *o, double *jj, double *rho) {
    const int GR = o->GR, zdf = o->zdf, zdc = o->zdc, NZ2 = orc_profile_len(o);
    int i, j, k, a;
    memset(jj, 0, sizeof(double) * NZ2);
    memset(rho, 0, sizeof(double) * NZ2);
#define ACC(fp, slot) do { double d_ = 0., x_ = 0.; \
        for (a = 0; a < 19; a++) { d_ += (fp)[a]; x_ += E0[a] * (fp)[a]; } \
        jj[slot] += x_ / d_; rho[slot] += d_; } while (0)
    for (i = 0; i < o->Xf; i++)                                /* :62-77 */
        for (j = 0; j < o->Yf; j++)
            for (k = 1; k < (zdf - GR - 1); k++) ACC(o->lft + FI(o, i, j, k), k);
    for (k = 1; k < (zdf - GR - 1); k++) {
        jj[k] /= (double)((o->Xf + 1 - 1.) * 1 * (o->Yf + 1 - 1.) * 1);
        rho[k] /= (double)((o->Xf + 1 - 1.) * 1 * (o->Yf + 1 - 1.) * 1);
    }
    for (i = 0; i < o->Xc; i++)                                /* :84-100 */
        for (j = 0; j < o->Yc; j++)
            for (k = 0; k < zdc; k++) ACC(o->cft + CI(o, i, j, k), zdf - GR - 1 + k);
    for (k = 0; k < zdc; k++) {
        jj[zdf - GR - 1 + k] /= (double)((o->Xc + 1 - 1.) * 1 * (o->Yc + 1 - 1.) * 1);
        rho[zdf - GR - 1 + k] /= (double)((o->Xc + 1 - 1.) * 1 * (o->Yc + 1 - 1.) * 1);
    }
    for (i = 0; i < o->Xf; i++)                                /* :108-124 */
        for (j = 0; j < o->Yf; j++)
            for (k = GR + 1; k < (zdf - 1); k++) ACC(o->uft + FI(o, i, j, k), zdf + zdc - GR - 2 - GR + k);
    for (k = GR + 1; k < (zdf - 1); k++) {
        jj[zdf + zdc - GR - 2 - GR + k] /= (double)((o->Xf + 1 - 1.) * 1 * (o->Yf + 1 - 1.) * 1);
        rho[zdf + zdc - GR - 2 - GR + k] /= (double)((o->Xf + 1 - 1.) * 1 * (o->Yf + 1 - 1.) * 1);
    }
#undef ACC
}

/* ---- one coarse step: main:1256-1375 (file I/O excluded) ---------------------------- */
static void step_after_reductions(orc_t *o) {
    int iter;
    orc_xcomputations(o);                   /* :1345 */
    for (iter = 1; iter < o->GR; iter++) {  /* :1347-1358 */
        orc_xfcomputations(o);
        orc_tinterpolate_hrmt(o, iter);
        orc_derivative(o, o->uint_);
        orc_derivative(o, o->lint);
        orc_ctfdtransfer(o, iter);
        orc_fexchange(o);
    }
    orc_xfcomputations(o);                  /* :1359-1365 */
    orc_tinterpolate_hrmt(o, o->GR);
    orc_derivative(o, o->uint_);
    orc_derivative(o, o->lint);
    orc_ctfdtransfer(o, o->GR);
    orc_ftcdtransfer(o);                    /* :1366-1368 */
    orc_exchange(o);                        /* :1373-1374 */
    orc_fexchange(o);
}

int orc_step(orc_t *o, long ss, dnsg_diag *dg) {
    int rc;
    orc_encalc2(o, ss, dg);                 /* :1257 */
    rc = orc_wshearstr(o, ss, dg);          /* :1258 */
    if (rc) return rc;
    step_after_reductions(o);
    return 0;
}

/* The same step with the body force that encalc2 forms replaced by g.  Test hook: the
   constant-force mode given, step by step, the force the flux controller formed must
   reproduce the controller's run, so that the two modes differ only in where the force
   comes from (encalc2.c:129 against a constant). */
int orc_step_forced(orc_t *o, long ss, dnsg_diag *dg, double g) {
    int rc;
    orc_encalc2(o, ss, dg);
    o->force[0] = g;
    if (dg) dg->force = g;
    rc = orc_wshearstr(o, ss, dg);
    if (rc) return rc;
    step_after_reductions(o);
    return 0;
}

/* ---- files in the reference layouts ------------------------------------------------ */
static int write_two(const char *path, const double *a, const double *b, size_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) return DNSG_ERR_IO;
    if (fwrite(a, 8, n, f) != n || (b && fwrite(b, 8, n, f) != n)) { fclose(f); return DNSG_ERR_IO; }
    return fclose(f) ? DNSG_ERR_IO : 0;
}

int orc_write_checkpoint(const orc_t *o, const char *dir, long index) {
    char fn[4096];
    size_t np = (size_t)o->Xc * o->Yc * 23;
    FILE *f;
    int rc = 0;
    snprintf(fn, sizeof fn, "%s/lbf-c.%.4ld", dir, index);   rc |= write_two(fn, o->cft, o->cf, o->d.n_coarse);
    snprintf(fn, sizeof fn, "%s/lbf-slf.%.4ld", dir, index); rc |= write_two(fn, o->lft, o->lf, o->d.n_band);
    snprintf(fn, sizeof fn, "%s/lbf-suf.%.4ld", dir, index); rc |= write_two(fn, o->uft, o->uf, o->d.n_band);
    snprintf(fn, sizeof fn, "%s/force.%.4ld", dir, index);   rc |= write_two(fn, &o->force[0], NULL, 1);
    snprintf(fn, sizeof fn, "%s/lbf-uin.%.4ld", dir, index);
    f = fopen(fn, "wb");
    if (!f) return DNSG_ERR_IO;
    fwrite(o->uin, 8, np, f); fwrite(o->uinp, 8, np, f); fwrite(o->uinpp, 8, np, f);
    fclose(f);
    snprintf(fn, sizeof fn, "%s/lbf-lin.%.4ld", dir, index);
    f = fopen(fn, "wb");
    if (!f) return DNSG_ERR_IO;
    fwrite(o->lin, 8, np, f); fwrite(o->linp, 8, np, f); fwrite(o->linpp, 8, np, f);
    fclose(f);
    return rc ? DNSG_ERR_IO : 0;
}

static int read_n(FILE *f, double *dst, size_t n) { return fread(dst, 8, n, f) == n ? 0 : -1; }

int orc_read_checkpoint(orc_t *o, const char *dir, long index) {
    char fn[4096];
    size_t np = (size_t)o->Xc * o->Yc * 23;
    FILE *f;
    int bad = 0;
    snprintf(fn, sizeof fn, "%s/lbf-c.%.4ld", dir, index);
    if (!(f = fopen(fn, "rb"))) return DNSG_ERR_IO;
    bad |= read_n(f, o->cft, o->d.n_coarse); bad |= read_n(f, o->cf, o->d.n_coarse); fclose(f);
    snprintf(fn, sizeof fn, "%s/lbf-slf.%.4ld", dir, index);
    if (!(f = fopen(fn, "rb"))) return DNSG_ERR_IO;
    bad |= read_n(f, o->lft, o->d.n_band); bad |= read_n(f, o->lf, o->d.n_band); fclose(f);
    snprintf(fn, sizeof fn, "%s/lbf-suf.%.4ld", dir, index);
    if (!(f = fopen(fn, "rb"))) return DNSG_ERR_IO;
    bad |= read_n(f, o->uft, o->d.n_band); bad |= read_n(f, o->uf, o->d.n_band); fclose(f);
    snprintf(fn, sizeof fn, "%s/force.%.4ld", dir, index);
    if (!(f = fopen(fn, "rb"))) return DNSG_ERR_IO;
    bad |= read_n(f, &o->force[0], 1); fclose(f);
    snprintf(fn, sizeof fn, "%s/lbf-uin.%.4ld", dir, index);
    if (!(f = fopen(fn, "rb"))) return DNSG_ERR_IO;
    bad |= read_n(f, o->uin, np); bad |= read_n(f, o->uinp, np); bad |= read_n(f, o->uinpp, np); fclose(f);
    snprintf(fn, sizeof fn, "%s/lbf-lin.%.4ld", dir, index);
    if (!(f = fopen(fn, "rb"))) return DNSG_ERR_IO;
    bad |= read_n(f, o->lin, np); bad |= read_n(f, o->linp, np); bad |= read_n(f, o->linpp, np); fclose(f);
    if (o->p.force_mode == DNSG_FORCE_CONST) o->force[0] = o->p.g_const;
    return bad ? DNSG_ERR_IO : 0;
}

static int append_line(const char *dir, const char *name, double t, double v) {
    char fn[4096];
    FILE *f;
    snprintf(fn, sizeof fn, "%s/%s", dir, name);
    if (!(f = fopen(fn, "a"))) return DNSG_ERR_IO;
    fprintf(f, "%f %.12f\n", t, v);
    fclose(f);
    return 0;
}

int orc_append_logs(const char *dir, const dnsg_diag *d) {
    int rc = 0;
    rc |= append_line(dir, "flowrate.txt", d->time, d->flowrate); /* encalc2.c:137-139 */
    rc |= append_line(dir, "density.txt", d->time, d->density);   /* :141-143 */
    rc |= append_line(dir, "str-up.txt", d->time, d->str_up);     /* wshearstr.c:146-149 */
    rc |= append_line(dir, "str-do.txt", d->time, d->str_down);   /* :298-301 */
    rc |= append_line(dir, "w-str.txt", d->time, d->stress);      /* :303-306 */
    rc |= append_line(dir, "p-grad.txt", d->time, d->pgrad);      /* :308-311 */
    return rc;
}

int orc_write_turbfld(const char *dir, long ts, const double *stats, int NZ) {
    char fn[4096];
    FILE *f;
    snprintf(fn, sizeof fn, "%s/turb-fld.%.3d.%.3ld", dir, 0, ts); /* statistics.c:305 */
    if (!(f = fopen(fn, "wb"))) return DNSG_ERR_IO;
    fwrite(stats, 8, (size_t)9 * NZ, f);
    fclose(f);
    return 0;
}

/* the driver loop of main:1254-1375 with checkpoints every 50*utime steps,
   statistics every utime steps and the per-step logs (vel_print omitted) */
int orc_run(orc_t *o, const char *dir, long up, int write_files) {
    long ss;
    const long ut = o->d.utime;
    double *stats = malloc(sizeof(double) * 9 * o->NZ);
    int rc = 0;
    for (ss = 0; ss < (up + 1); ss++) {
        dnsg_diag d;
        long TM = o->p.TSTR;
        memset(&d, 0, sizeof d);
        orc_encalc2(o, ss, &d);
        rc = orc_wshearstr(o, ss, &d);
        if (write_files) orc_append_logs(dir, &d);
        if (rc) break;
        if (!(ss % ut)) {
            if (!(ss % (50 * ut)) && write_files) orc_write_checkpoint(o, dir, TM + (ss / ut));
            if (write_files) orc_vel_print(o, dir, TM + (ss / ut));
            orc_statistics(o, stats);
            if (write_files) orc_write_turbfld(dir, TM + (ss / ut), stats, o->NZ);
        }
        {
            int iter;
            orc_xcomputations(o);
            for (iter = 1; iter < o->GR; iter++) {
                orc_xfcomputations(o);
                orc_tinterpolate_hrmt(o, iter);
                orc_derivative(o, o->uint_);
                orc_derivative(o, o->lint);
                orc_ctfdtransfer(o, iter);
                orc_fexchange(o);
            }
            orc_xfcomputations(o);
            orc_tinterpolate_hrmt(o, o->GR);
            orc_derivative(o, o->uint_);
            orc_derivative(o, o->lint);
            orc_ctfdtransfer(o, o->GR);
            orc_ftcdtransfer(o);
            orc_exchange(o);
            orc_fexchange(o);
        }
    }
    if (!rc && write_files) { /* output_twoD2 (main:1379) and the final statistics (main:1414) */
        int n2 = orc_profile_len(o);
        double *jj = malloc(sizeof(double) * n2), *rh = malloc(sizeof(double) * n2);
        char fn[4096];
        FILE *f;
        orc_profile2d(o, jj, rh);
        snprintf(fn, sizeof fn, "%s/2dresults-lb.bdat", dir);
        if ((f = fopen(fn, "wb"))) { fwrite(jj, 8, n2, f); fclose(f); }
        free(jj); free(rh);
        orc_statistics(o, stats);
        orc_write_turbfld(dir, o->p.TSTR + (ss / ut), stats, o->NZ);
    }
    free(stats);
    return rc;
}
