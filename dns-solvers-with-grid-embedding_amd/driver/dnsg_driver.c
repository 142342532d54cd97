/*
 * dnsg_driver.c -- the kept C driver: the reference's main() (dns-q-mgrd-all-
 * patterns-xyc.c:15-1535) on one GPU through the dnsg_* C-ABI.
 *
 *   dnsg-run -s|-p [definitions.h] [key=value ...]
 *
 * Parameters come from a definitions.h-format file (the reference's own
 * `#define NAME value` lines are parsed, including the RSLIP/PSLIP/SSLIP and
 * ITER switches and the NOFLTR/FLTR injection choice) and/or key=value
 * arguments with the dnsg_params field names, which also reach fltr and the
 * runtime-only extensions (force_mode, g_const, tau_c, precision, init_kind,
 * seed, u_max, noise, results_dat).  print_params=1 prints
 * the parsed parameters and the derived sizes and exits without GPU work;
 * spectra=1 also writes the 1-D energy spectra of every velocity snapshot (the
 * files of the reference's e-spectra tool, computed on the device); budget=1 the
 * per-snapshot statistics and energy budget of its TKE-budget tool.
 *
 * Run modes (main:944-1190): -i interpolates the uniform-grid file lbf.TSTR
 * (sizes xdu, ydu, zdu) onto the grids, as the reference's '-i' does;
 * -s starts from rest ('-s' of the reference) or,
 * with init_kind=1, from the seeded laminar profile with noise; -p reads the
 * restart files lbf-c/slf/suf/lin/uin.TSTR and force.TSTR from the working
 * directory (their sizes are validated, which the reference does not do).
 *
 * Output files, written to the working directory in the reference's names and
 * layouts: flowrate.txt, density.txt, str-up.txt, str-do.txt, w-str.txt,
 * p-grad.txt (every step); vel-{c,slf,suf}.NNNN, den-*.NNNN, turb-fld.000.NNN,
 * uavg-temp.dat, uavg-exact-lam.dat (every utime steps); lbf-*.NNNN and
 * force.NNNN (every 50 utime steps); 2dresults-exact.dat, 2dresults-lb.dat,
 * 2dresults-lb.bdat, report.code and results.dat at the end.
 */
#define _POSIX_C_SOURCE 200809L
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/dnsg.h"

static const double E[3][19] = {{0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0},
                                {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1},
                                {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1}};

static int results_dat = 1;
static int uni_size[3] = {0, 0, 0}; /* xdu, ydu, zdu of the '-i' uniform grid (definitions.h:107-109) */
static int print_params = 0;
static int spectra_out = 0;
static int budget_out = 0;

static double now(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void die(dnsg_t *h, const char *what, int rc) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, h ? dnsg_last_error(h) : "");
    exit(1);
}

/* one NAME=value assignment (definitions.h #define or command line) */
static int set_param(dnsg_params *p, const char *key, const char *val) {
#define D(n) if (!strcmp(key, #n)) { p->n = strtod(val, NULL); return 0; }
#define I(n) if (!strcmp(key, #n)) { p->n = (int)strtol(val, NULL, 10); return 0; }
    D(Rebs) D(ub) I(zdv) I(ydv) I(xdv) I(TSTR) I(TEND) D(CFL) I(gratio) I(zds)
    I(gyplussize) I(wyplussize) I(gxplussize) I(wxplussize) I(pattern)
    I(force_mode) D(g_const) D(tau_c) I(precision) I(init_kind) D(u_max) D(noise) I(fltr)
#undef D
#undef I
    if (!strcmp(key, "ITER")) { p->ITER = strtol(val, NULL, 10); return 0; }
    if (!strcmp(key, "seed")) { p->seed = strtoull(val, NULL, 10); return 0; }
    if (!strcmp(key, "results_dat")) { results_dat = atoi(val); return 0; }
    if (!strcmp(key, "print_params")) { print_params = atoi(val); return 0; }
    if (!strcmp(key, "spectra")) { spectra_out = atoi(val); return 0; }
    if (!strcmp(key, "budget")) { budget_out = atoi(val); return 0; }
    if (!strcmp(key, "XDIM") || !strcmp(key, "YDIM")) return 0; /* one GPU = one rank */
    if (!strcmp(key, "xdu")) { uni_size[0] = atoi(val); return 0; }
    if (!strcmp(key, "ydu")) { uni_size[1] = atoi(val); return 0; }
    if (!strcmp(key, "zdu")) { uni_size[2] = atoi(val); return 0; }
    return -1;
}

/* parse a definitions.h-format file: active `#define NAME value` lines */
static int read_definitions(dnsg_params *p, const char *path) {
    FILE *f = fopen(path, "r");
    char line[512];
    if (!f) return -1;
    p->pattern = DNSG_PATTERN_NONE;
    p->ITER = 0;
    p->fltr = 0;
    while (fgets(line, sizeof line, f)) {
        char name[128], val[256];
        const char *s = line;
        while (isspace((unsigned char)*s)) s++;
        if (strncmp(s, "#define", 7)) continue;
        val[0] = 0;
        if (sscanf(s + 7, "%127s %255s", name, val) < 1) continue;
        if (!strcmp(name, "RSLIP")) p->pattern = DNSG_PATTERN_RSLIP;
        else if (!strcmp(name, "PSLIP")) p->pattern = DNSG_PATTERN_PSLIP;
        else if (!strcmp(name, "SSLIP")) p->pattern = DNSG_PATTERN_SSLIP;
        else if (!strcmp(name, "FLTR")) p->fltr = 1;   /* main:1369-1371 */
        else if (!strcmp(name, "NOFLTR")) continue;     /* main:1366-1368, the default */
        else if (val[0]) set_param(p, name, val);
    }
    fclose(f);
    return 0;
}

static int read_doubles(const char *fn, double *dst, size_t n, size_t expect_total) {
    FILE *f = fopen(fn, "rb");
    struct stat st;
    if (!f) { fprintf(stderr, "cannot open %s\n", fn); return -1; }
    if (fstat(fileno(f), &st) || (size_t)st.st_size != expect_total * 8) {
        fprintf(stderr, "%s: size %ld, expected %zu bytes\n", fn, (long)st.st_size, expect_total * 8);
        fclose(f);
        return -1;
    }
    if (fread(dst, 8, n, f) != n) { fclose(f); return -1; }
    fclose(f);
    return 0;
}

static int write_file(const char *fn, const double *a, size_t n, const double *b, size_t m) {
    FILE *f = fopen(fn, "wb");
    if (!f) return -1;
    fwrite(a, 8, n, f);
    if (b) fwrite(b, 8, m, f);
    return fclose(f);
}

static void append_log(const char *fn, double t, double v) {
    FILE *f = fopen(fn, "a");
    if (!f) return;
    fprintf(f, "%f %.12f\n", t, v);
    fclose(f);
}

/* checkpoint (main:1262-1341): ftemp then f for every lattice; the second half
 * of the reference's files is a scratch buffer that the restart never reads, it
 * holds the state again here */
static void checkpoint(dnsg_t *h, const dnsg_dims *d, long idx, double force) {
    double *c = malloc(d->n_coarse * 8), *lo = malloc(d->n_band * 8), *up = malloc(d->n_band * 8);
    double *li = malloc(d->n_plane * 8), *ui = malloc(d->n_plane * 8);
    char fn[64];
    int rc;
    if (!c || !lo || !up || !li || !ui) { fprintf(stderr, "checkpoint: out of memory\n"); exit(1); }
    if ((rc = dnsg_download_state(h, c, lo, up, li, ui, NULL))) die(h, "dnsg_download_state", rc);
    snprintf(fn, sizeof fn, "lbf-c.%.4ld", idx);   write_file(fn, c, d->n_coarse, c, d->n_coarse);
    snprintf(fn, sizeof fn, "lbf-slf.%.4ld", idx); write_file(fn, lo, d->n_band, lo, d->n_band);
    snprintf(fn, sizeof fn, "lbf-suf.%.4ld", idx); write_file(fn, up, d->n_band, up, d->n_band);
    snprintf(fn, sizeof fn, "force.%.4ld", idx);   write_file(fn, &force, 1, NULL, 0);
    snprintf(fn, sizeof fn, "lbf-uin.%.4ld", idx); write_file(fn, ui, d->n_plane, NULL, 0);
    snprintf(fn, sizeof fn, "lbf-lin.%.4ld", idx); write_file(fn, li, d->n_plane, NULL, 0);
    printf("writing V.force=%.12f num written=1 myrank=0\n", force);
    free(c); free(lo); free(up); free(li); free(ui);
}

/* vel_print.c:15-274 */
static void vel_print(dnsg_t *h, const dnsg_dims *d, long cntr) {
    static const char *tag[3] = {"c", "slf", "suf"};
    static const int order[3] = {0, 2, 1};
    int gi;
    for (gi = 0; gi < 3; gi++) {
        const int g = order[gi];
        const size_t n = g == 0 ? (size_t)d->Xc * d->Yc * d->zdc : (size_t)d->Xf * d->Yf * d->zdf;
        double *vel = malloc(n * 3 * 8), *den = malloc(n * 8);
        char fn[64];
        int rc;
        if (!vel || !den) { fprintf(stderr, "vel_print: out of memory\n"); exit(1); }
        if ((rc = dnsg_velocity_snapshot(h, g, vel, den))) die(h, "dnsg_velocity_snapshot", rc);
        snprintf(fn, sizeof fn, "vel-%s.%.4d", tag[g], (int)cntr); write_file(fn, vel, n * 3, NULL, 0);
        snprintf(fn, sizeof fn, "den-%s.%.4d", tag[g], (int)cntr); write_file(fn, den, n, NULL, 0);
        free(vel); free(den);
    }
}

/* e-spectra/src/1d-spectra-mgrd.c:228-265 file part: the 1-D spectra of the three
   grids from the current state, files E{uu,vv,ww}{x,y}-{c,fl,fu}.NNNN */
static void spectra_print(dnsg_t *h, const dnsg_dims *d, long cntr) {
    static const char *tag[3] = {"c", "fl", "fu"};
    static const char *comp[3] = {"uu", "vv", "ww"};
    int g, c;
    for (g = 0; g < 3; g++) {
        const int N0 = g == 0 ? d->Xc : d->Xf, N1h = (g == 0 ? d->Yc : d->Yf) / 2 + 1;
        const int nz = g == 0 ? d->zdc : d->zdf;
        double *ex = malloc(sizeof(double) * 3 * N0 * nz), *ey = malloc(sizeof(double) * 3 * N1h * nz);
        char fn[64];
        int rc;
        if (!ex || !ey) { fprintf(stderr, "spectra: out of memory\n"); exit(1); }
        if ((rc = dnsg_spectra(h, g, ex, ey))) die(h, "dnsg_spectra", rc);
        for (c = 0; c < 3; c++) {
            snprintf(fn, sizeof fn, "E%sx-%s.%.4d", comp[c], tag[g], (int)cntr);
            write_file(fn, ex + (size_t)c * N0 * nz, (size_t)N0 * nz, NULL, 0);
            snprintf(fn, sizeof fn, "E%sy-%s.%.4d", comp[c], tag[g], (int)cntr);
            write_file(fn, ey + (size_t)c * N1h * nz, (size_t)N1h * nz, NULL, 0);
        }
        free(ex); free(ey);
    }
}

/* TKE-budget/computations/budget2D.c:1550-1636 and uv_fluc.c:35-45 file part: the
   tool's per-snapshot files turb-fld-2d, vort-fluc-2d, KE-Budget-2d.000.NNNN and,
   every 200th snapshot, ufl4/vfl4.NNNN, from the current state */
static void budget_print(dnsg_t *h, const dnsg_dims *d, long cntr) {
    const int NZ = d->NZ, dump = !(cntr % 200);
    const size_t np = (size_t)d->Xf * d->Yf;
    double *turb = malloc(sizeof(double) * 19 * NZ), *vort = malloc(sizeof(double) * 3 * NZ);
    double *ke = malloc(sizeof(double) * 30 * NZ);
    double *ufl = dump ? malloc(sizeof(double) * np) : NULL, *vfl = dump ? malloc(sizeof(double) * np) : NULL;
    char fn[64];
    int rc;
    if (!turb || !vort || !ke || (dump && (!ufl || !vfl))) { fprintf(stderr, "budget: out of memory\n"); exit(1); }
    if ((rc = dnsg_budget2d(h, turb, vort, ke, ufl, vfl))) die(h, "dnsg_budget2d", rc);
    snprintf(fn, sizeof fn, "turb-fld-2d.%.3d.%.4d", 0, (int)cntr);
    write_file(fn, turb, (size_t)19 * NZ, NULL, 0);
    snprintf(fn, sizeof fn, "vort-fluc-2d.%.3d.%.4d", 0, (int)cntr);
    write_file(fn, vort, (size_t)3 * NZ, NULL, 0);
    snprintf(fn, sizeof fn, "KE-Budget-2d.%.3d.%.4d", 0, (int)cntr);
    write_file(fn, ke, (size_t)30 * NZ, NULL, 0);
    if (dump) {
        snprintf(fn, sizeof fn, "ufl4.%.4d", (int)cntr);
        write_file(fn, ufl, np, NULL, 0);
        snprintf(fn, sizeof fn, "vfl4.%.4d", (int)cntr);
        write_file(fn, vfl, np, NULL, 0);
    }
    free(turb); free(vort); free(ke); free(ufl); free(vfl);
}

/* wall-normal coordinates of the profile slots (statistics.c:319-329) */
static void profile_coords(const dnsg_dims *d, double *zps) {
    const int zdf = d->zdf, zdc = d->zdc, GR = d->GR, NZ = d->NZ;
    int k;
    memset(zps, 0, sizeof(double) * NZ);
    for (k = 1; k < zdf; k++) zps[k] = k - 0.5;
    for (k = 2; k < (zdc - 2); k++) zps[zdf - 2 + k] = GR * (k - 1) + (zdf - 1 - 0.5);
    for (k = 0; k < (zdf - 1); k++) zps[zdf + zdc - 4 + k] = k + GR * (zdc - 3) + zdf - 1. - 0.5;
    zps[2 * zdf + zdc - 5] = 2 * zdf + GR * (zdc - 4) + GR - 3;
    for (k = 0; k < NZ; k++) zps[k] /= ((double)(d->NZs - 2.));
}

/* statistics.c:303-357 file part */
static void statistics(dnsg_t *h, const dnsg_dims *d, const dnsg_params *p, long ts) {
    const int NZ = d->NZ, NZs = d->NZs;
    double *st = malloc(sizeof(double) * 9 * NZ), *zps = malloc(sizeof(double) * NZ);
    double *y = calloc(NZs, sizeof(double)), *zpsf = calloc(NZs, sizeof(double));
    char fn[64];
    FILE *f;
    int k, rc;
    if ((rc = dnsg_statistics(h, st))) die(h, "dnsg_statistics", rc);
    snprintf(fn, sizeof fn, "turb-fld.%.3d.%.3ld", 0, ts);
    write_file(fn, st, (size_t)9 * NZ, NULL, 0);
    profile_coords(d, zps);
    for (k = 1; k < (NZs - 1); k++)
        y[k] = -6. * p->ub * ((k - 0.5) / (NZs - 2.)) * ((k - 0.5) / (NZs - 2.) - 1);
    if ((f = fopen("uavg-temp.dat", "w"))) {
        for (k = 0; k < NZ; k++) fprintf(f, "%.12f %.12f\n", zps[k], st[k]);
        fclose(f);
    }
    for (k = 1; k < (NZs - 1); k++) zpsf[k] = k - 0.5;
    zpsf[NZs - 1] = NZs - 2;
    for (k = 0; k < NZs; k++) zpsf[k] /= ((double)(NZs - 2.));
    if ((f = fopen("uavg-exact-lam.dat", "w"))) {
        for (k = 0; k < NZs; k++) fprintf(f, "%.12f %.12f \n", zpsf[k], y[k]);
        fclose(f);
    }
    free(st); free(zps); free(y); free(zpsf);
}

static double node_ux(const double *f, double *dens) {
    double ux = 0., d = 0.;
    int a;
    for (a = 0; a < 19; a++) { d += f[a]; ux += E[0][a] * f[a]; }
    *dens = d;
    return ux;
}

/* output_twoD2.c:15-169 (the files it writes replace output_twoD's) */
static void output_twoD2(const dnsg_dims *d, const dnsg_params *p, const double *c, const double *lo,
                         const double *up) {
    const int GR = d->GR, zdf = d->zdf, zdc = d->zdc, NZs = d->NZs, Xf = d->Xf, Yf = d->Yf, Xc = d->Xc, Yc = d->Yc;
    const int NZ = 2 * (zdf - GR - 1) + zdc;
    double *jj = calloc(NZ, 8), *zps = calloc(NZ, 8), *y = calloc(NZs, 8), *zpsf = calloc(NZs, 8);
    int i, j, k;
    FILE *f;
    for (k = 1; k < (NZs - 1); k++) zpsf[k] = k - 0.5;
    zpsf[NZs - 1] = NZs - 2;
    for (k = 0; k < NZs; k++) zpsf[k] /= ((double)(NZs - 2.));
    for (k = 1; k < (zdf - GR - 1); k++) zps[k] = k - 0.5;
    for (k = 0; k < zdc; k++) zps[zdf - GR - 1 + k] = GR * (k) + (zdf - GR - 1 - 0.5);
    for (k = GR + 1; k < (zdf - 1); k++) zps[zdf + zdc - GR - 2 - GR + k] = k - GR + GR * (zdc - 1) + zdf - GR - 1. - 0.5;
    zps[NZ - 1] = 2 * zdf + GR * (zdc - 4) + GR - 3;
    for (k = 0; k < NZ; k++) zps[k] /= ((double)(NZs - 2.));
    for (k = 1; k < (NZs - 1); k++) y[k] = -6. * p->ub * ((k - 0.5) / (NZs - 2.)) * ((k - 0.5) / (NZs - 2.) - 1);
    for (i = 0; i < Xf; i++)
        for (j = 0; j < Yf; j++)
            for (k = 1; k < (zdf - GR - 1); k++) {
                double dens, ux = node_ux(lo + (((size_t)i * Yf + j) * zdf + k) * 19, &dens);
                jj[k] += ux / dens;
            }
    for (k = 1; k < (zdf - GR - 1); k++) jj[k] /= (double)((Xf + 0.) * 1 * (Yf + 0.) * 1);
    for (i = 0; i < Xc; i++)
        for (j = 0; j < Yc; j++)
            for (k = 0; k < zdc; k++) {
                double dens, ux = node_ux(c + (((size_t)i * Yc + j) * zdc + k) * 19, &dens);
                jj[zdf - GR - 1 + k] += ux / dens;
            }
    for (k = 0; k < zdc; k++) jj[zdf - GR - 1 + k] /= (double)((Xc + 0.) * 1 * (Yc + 0.) * 1);
    for (i = 0; i < Xf; i++)
        for (j = 0; j < Yf; j++)
            for (k = GR + 1; k < (zdf - 1); k++) {
                double dens, ux = node_ux(up + (((size_t)i * Yf + j) * zdf + k) * 19, &dens);
                jj[zdf + zdc - GR - 2 - GR + k] += ux / dens;
            }
    for (k = GR + 1; k < (zdf - 1); k++) jj[zdf + zdc - GR - 2 - GR + k] /= (double)((Xf + 0.) * 1 * (Yf + 0.) * 1);
    if ((f = fopen("2dresults-exact.dat", "w"))) {
        for (k = 0; k < NZs; k++) fprintf(f, "%.12f %.12f \n", zpsf[k], y[k]);
        fclose(f);
    }
    if ((f = fopen("2dresults-lb.dat", "w"))) {
        for (k = 0; k < NZ; k++) fprintf(f, "%.12f %.12f\n", zps[k], jj[k]);
        fclose(f);
    }
    write_file("2dresults-lb.bdat", jj, NZ, NULL, 0);
    free(jj); free(zps); free(y); free(zpsf);
}

/* output.c:15-107 (one rank): Tecplot points of the three grids */
static void output_tecplot(const dnsg_dims *d, const dnsg_params *p, const double *c, const double *lo,
                           const double *up) {
    const int GR = d->GR, zdf = d->zdf, zdc = d->zdc, Xf = d->Xf, Yf = d->Yf, Xc = d->Xc, Yc = d->Yc;
    const double h = (d->NZs - 2.) / 2., ub = p->ub;
    double *zps = malloc(sizeof(double) * d->NZ);
    FILE *f = fopen("results.dat", "w");
    int i, j, k, a;
    if (!f) { free(zps); return; }
    {   /* unnormalised slot coordinates (output.c:25-33) */
        dnsg_dims t = *d;
        t.NZs = 3; /* profile_coords divides by NZs-2: keep the raw values */
        profile_coords(&t, zps);
    }
    fprintf(f, "TITLE = \"3d velocity field\"\n");
    fprintf(f, "VARIABLES = \"x/h\", \"y/h\", \"z/h\", \"rho\", \"u\", \"v\", \"w\" \n");
    fprintf(f, "ZONE I=%d, J=%d, K=%d, F=POINT", Xf, Yf + 1, zdf);
#define NODE(arr, Y, Z, ii, jj, kk)                                                                    \
    do {                                                                                               \
        const double *fp = (arr) + (((size_t)(ii) * (Y) + (jj)) * (Z) + (kk)) * 19;                    \
        ux = 0.; uy = 0.; uz = 0.; dens = 0.;                                                          \
        for (a = 0; a < 19; a++) { ux += E[0][a] * fp[a]; uy += E[1][a] * fp[a]; uz += E[2][a] * fp[a]; dens += fp[a]; } \
    } while (0)
    for (k = 0; k < zdf; k++)
        for (j = 0; j < Yf + 1; j++)
            for (i = 0; i < Xf; i++) {
                double ux, uy, uz, dens;
                NODE(lo, Yf, zdf, i, j % Yf, k);
                fprintf(f, "%f %f %f %f %f %f %f\n", (i + 0.5) / h, (j + 0.5) / h, (k - 0.5) / h, dens, ux / ub, uy / ub, uz / ub);
            }
    fprintf(f, "ZONE I=%d, J=%d, K=%d, F=POINT", Xc, Yc + 1, zdc - 4);
    for (k = 2; k < (zdc - 2); k++)
        for (j = 0; j < Yc + 1; j++)
            for (i = 0; i < Xc; i++) {
                double ux, uy, uz, dens;
                NODE(c, Yc, zdc, i, j % Yc, k);
                fprintf(f, "%f %f %f %f %f %f %f\n", (i * GR + 0.5) / h, (j * GR + 0.5) / h, zps[zdf - 2 + k] / h, dens, ux / ub, uy / ub, uz / ub);
            }
    fprintf(f, "ZONE I=%d, J=%d, K=%d, F=POINT", Xf, Yf + 1, zdf);
    for (k = 0; k < zdf; k++)
        for (j = 0; j < Yf + 1; j++)
            for (i = 0; i < Xf; i++) {
                double ux, uy, uz, dens;
                NODE(up, Yf, zdf, i, j % Yf, k);
                fprintf(f, "%f %f %f %f %f %f %f\n", (i + 0.5) / h, (j + 0.5) / h, zps[zdf + zdc - 4 + k] / h, dens, ux / ub, uy / ub, uz / ub);
            }
#undef NODE
    fprintf(f, "\n\n");
    fclose(f);
    free(zps);
}

int main(int argc, char **argv) {
    dnsg_params p;
    dnsg_dims d;
    dnsg_t *h = NULL;
    int i, rc;
    char mode;
    long ss, TM, up, ut;
    double t4, t2, mass;
    if (argc < 2 || argv[1][0] != '-' || !argv[1][1]) {
        fprintf(stderr, "usage: %s -s|-p|-i [definitions.h] [key=value ...]\n", argv[0]);
        return 2;
    }
    mode = argv[1][1];
    dnsg_params_default(&p);
    for (i = 2; i < argc; i++) {
        const char *eq = strchr(argv[i], '=');
        if (!eq) {
            if (read_definitions(&p, argv[i])) { fprintf(stderr, "cannot read %s\n", argv[i]); return 2; }
        } else {
            char key[128];
            size_t n = (size_t)(eq - argv[i]);
            if (n >= sizeof key) return 2;
            memcpy(key, argv[i], n);
            key[n] = 0;
            if (set_param(&p, key, eq + 1)) { fprintf(stderr, "unknown parameter %s\n", key); return 2; }
        }
    }
    if (print_params) { /* the parsed parameters and derived sizes; no GPU work */
        if ((rc = dnsg_dims_from_params(&p, &d))) { fprintf(stderr, "invalid parameters (%d)\n", rc); return 2; }
        printf("Rebs=%.17g\nub=%.17g\nxdv=%d\nydv=%d\nzdv=%d\nTSTR=%d\nTEND=%d\nCFL=%.17g\ngratio=%d\nzds=%d\n",
               p.Rebs, p.ub, p.xdv, p.ydv, p.zdv, p.TSTR, p.TEND, p.CFL, p.gratio, p.zds);
        printf("gyplussize=%d\nwyplussize=%d\ngxplussize=%d\nwxplussize=%d\npattern=%d\nITER=%ld\n",
               p.gyplussize, p.wyplussize, p.gxplussize, p.wxplussize, p.pattern, p.ITER);
        printf("fltr=%d\n", p.fltr);
        printf("force_mode=%d\nprecision=%d\ninit_kind=%d\nNZs=%d\nNZ=%d\nutime=%ld\nup=%ld\ntau_c=%.17g\ntau_f=%.17g\n",
               p.force_mode, p.precision, p.init_kind, d.NZs, d.NZ, d.utime, d.up, d.tau_c, d.tau_f);
        return 0;
    }
    if ((rc = dnsg_create(&p, &h))) die(h, "dnsg_create", rc);
    dnsg_get_dims(h, &d);
    printf("%s: is running the process 0 of total 1 processors\n", "dnsg-run");
    switch (mode) {
    case 's':
        TM = p.TSTR;
        if ((rc = dnsg_init_state(h))) die(h, "dnsg_init_state", rc);
        break;
    case 'p': {
        char fn[64];
        double *c = malloc(d.n_coarse * 8), *lo = malloc(d.n_band * 8), *upb = malloc(d.n_band * 8);
        double *li = malloc(d.n_plane * 8), *ui = malloc(d.n_plane * 8), force = 0.;
        const long cntr = p.TSTR;
        fprintf(stderr, "TRYING TO READ DATA FROM FILES\n");
        if (!c || !lo || !upb || !li || !ui) { fprintf(stderr, "out of memory\n"); return 1; }
        snprintf(fn, sizeof fn, "lbf-c.%.4ld", cntr);
        if (read_doubles(fn, c, d.n_coarse, 2 * d.n_coarse)) return 1;
        snprintf(fn, sizeof fn, "lbf-slf.%.4ld", cntr);
        if (read_doubles(fn, lo, d.n_band, 2 * d.n_band)) return 1;
        snprintf(fn, sizeof fn, "lbf-suf.%.4ld", cntr);
        if (read_doubles(fn, upb, d.n_band, 2 * d.n_band)) return 1;
        snprintf(fn, sizeof fn, "force.%.4ld", cntr);
        if (read_doubles(fn, &force, 1, 1)) return 1;
        snprintf(fn, sizeof fn, "lbf-uin.%.4ld", cntr);
        if (read_doubles(fn, ui, d.n_plane, d.n_plane)) return 1;
        snprintf(fn, sizeof fn, "lbf-lin.%.4ld", cntr);
        if (read_doubles(fn, li, d.n_plane, d.n_plane)) return 1;
        printf("reading V.force=%.12f myrank=0\n", force);
        if ((rc = dnsg_upload_state(h, c, lo, upb, li, ui, force))) die(h, "dnsg_upload_state", rc);
        free(c); free(lo); free(upb); free(li); free(ui);
        fprintf(stderr, "END OF READING DATA FROM FILES\nSTARTING FROM %ld\n", cntr);
        TM = cntr;
        break;
    }
    case 'i': { /* main:982-997: interpolate the uniform-grid file lbf.TSTR */
        char fn[64];
        const size_t nu = (size_t)uni_size[0] * uni_size[1] * (uni_size[2] + 2) * 19;
        double *u = nu ? malloc(nu * 8) : NULL;
        if (!u) { fprintf(stderr, "'-i' needs the uniform grid sizes xdu, ydu, zdu\n"); return 2; }
        snprintf(fn, sizeof fn, "lbf.%.4d", p.TSTR); /* initializer.c:303 */
        fprintf(stderr, "starting to read and interpolate the data 0\n");
        if (read_doubles(fn, u, nu, 2 * nu)) return 1; /* the ftemp half */
        if ((rc = dnsg_init_uniform(h, u, uni_size[0], uni_size[1], uni_size[2]))) die(h, "dnsg_init_uniform", rc);
        free(u);
        TM = p.TSTR;
        break;
    }
    default:
        fprintf(stderr, "not a good file name\n");
        return 1;
    }
    if ((rc = dnsg_mass(h, &mass))) die(h, "dnsg_mass", rc);
    printf("Average Density in the channel at start: %.12f  myrank=0\n", mass);

    up = d.up;
    ut = d.utime;
    t4 = now();
    ss = 0;
    while (ss < up + 1) { /* main:1254-1375 */
        long next = up + 1, n, s;
        dnsg_diag *dg;
        if (!(ss % ut)) {
            /* output due at the top of this step, after encalc2/wshearstr */
            dnsg_diag d0;
            if ((rc = dnsg_encalc2(h, ss, &d0))) die(h, "dnsg_encalc2", rc);
            if (!(ss % (50 * ut))) checkpoint(h, &d, TM + (ss / ut), d0.force);
            vel_print(h, &d, TM + (ss / ut));
            if (spectra_out) spectra_print(h, &d, TM + (ss / ut));
            if (budget_out) budget_print(h, &d, TM + (ss / ut));
            statistics(h, &d, &p, TM + (ss / ut));
        }
        /* run up to the next output step (or the end) in one batch */
        next = ((ss / ut) + 1) * ut;
        if (next > up + 1) next = up + 1;
        n = next - ss;
        dg = calloc((size_t)n, sizeof(dnsg_diag));
        if (!dg) { fprintf(stderr, "out of memory\n"); return 1; }
        rc = dnsg_step(h, (int)n, dg);
        if (rc && rc != DNSG_ERR_NAN) { /* no record of an unfinished batch is logged */
            free(dg);
            die(h, "dnsg_step", rc);
        }
        /* on NaN the records end at the failing step, whose flowrate/density lines
           encalc2 has written before wshearstr aborts (wshearstr.c:139-143) */
        for (s = 0; s < n; s++) {
            if (rc == DNSG_ERR_NAN && isnan(dg[s].stress)) {
                append_log("flowrate.txt", dg[s].time, dg[s].flowrate);
                append_log("density.txt", dg[s].time, dg[s].density);
                fprintf(stderr, "Error: NAN in wal shear stress\n");
                return 1;
            }
            append_log("flowrate.txt", dg[s].time, dg[s].flowrate);
            append_log("density.txt", dg[s].time, dg[s].density);
            append_log("str-up.txt", dg[s].time, dg[s].str_up);
            append_log("str-do.txt", dg[s].time, dg[s].str_down);
            append_log("w-str.txt", dg[s].time, dg[s].stress);
            append_log("p-grad.txt", dg[s].time, dg[s].pgrad);
        }
        free(dg);
        if (rc) die(h, "dnsg_step", rc);
        ss = next;
    }
    t2 = now();
    {   /* end-of-run outputs (main:1377-1414) */
        double *c = malloc(d.n_coarse * 8), *lo = malloc(d.n_band * 8), *upb = malloc(d.n_band * 8);
        FILE *tm;
        if (!c || !lo || !upb) { fprintf(stderr, "out of memory\n"); return 1; }
        if ((rc = dnsg_download_state(h, c, lo, upb, NULL, NULL, NULL))) die(h, "dnsg_download_state", rc);
        output_twoD2(&d, &p, c, lo, upb);
        if ((tm = fopen("report.code", "w"))) {
            fprintf(tm, "xdc=%d, ydc=%d, zdc+2=%d, Reb=%f, tauc=%f, G=%.10f, Ub=%f, CFLnumber=%f\n", d.Xc + 1,
                    d.Yc + 1, d.zdc, p.Rebs, d.omega_c, 0., p.ub, p.CFL);
            fprintf(tm, "xdf=%d, ydf=%d, zdf+2=%d, Reb=%f, tauf=%f, G=%.10f, Ub=%f, CFLnumber=%f\n", d.Xf + 1,
                    d.Yf + 1, d.zdf, p.Rebs, d.omega_f, 0., p.ub, p.CFL);
            fprintf(tm, "NZ fine +2 = %d   GR=%d \n", d.NZs, d.GR);
            fprintf(tm, "number of processors: %d\n", 1);
            fprintf(tm, "Executed File's Name: %s\n", __FILE__);
            fprintf(tm, "Compilation Date: %s\n", __DATE__);
            fprintf(tm, "Compilation Time: %s\n", __TIME__);
            fprintf(tm, "Compiled wit standard C Compiler? (1 means yes): %d\n", __STDC__);
            fprintf(tm, "Compilation Option: xcut\n");
            if (p.pattern == DNSG_PATTERN_RSLIP) fprintf(tm, "slip option: RSLIP\n");
            if (p.pattern == DNSG_PATTERN_PSLIP) fprintf(tm, "slip option: PSLIP\n");
            if (p.pattern == DNSG_PATTERN_SSLIP) fprintf(tm, "slip option: SSLIP\n");
            fprintf(tm, "Total computational time is: %f s\n", t2 - t4);
            fprintf(tm, "Unit non-dimensional time is: %f iterations\n", (double)ut);
            fprintf(tm, "Starting time: %d\n", p.TSTR);
            fprintf(tm, "Current Time is: %d\n", p.TEND);
            fprintf(tm, "size of lbf-c: %lu lbf-f: %lu\n", (unsigned long)(d.n_coarse * 2 * 8), (unsigned long)(d.n_band * 2 * 8));
            fprintf(tm, "size of vel-c: %lu vel-f: %lu\n", (unsigned long)(d.n_coarse / 19 * 3 * 8), (unsigned long)(d.n_band / 19 * 3 * 8));
            fprintf(tm, "size of den-c: %lu den-f: %lu\n", (unsigned long)(d.n_coarse / 19 * 8), (unsigned long)(d.n_band / 19 * 8));
            fprintf(tm, "size of double: %lu\n", (unsigned long)sizeof(double));
            fprintf(tm, "size of uin %lu\n", (unsigned long)(d.n_plane * 8));
            fclose(tm);
        }
        if (results_dat) output_tecplot(&d, &p, c, lo, upb);
        free(c); free(lo); free(upb);
    }
    statistics(h, &d, &p, TM + (ss / ut));
    if ((rc = dnsg_mass(h, &mass))) die(h, "dnsg_mass", rc);
    printf("Average Density in the channel at end: %.12f  myrank=0\n", mass);
    printf("GLUPS: %.4f over %ld steps (%.3f s)\n",
           (double)(up + 1) * ((double)d.Xc * d.Yc * d.zdc + d.GR * 2. * d.Xf * d.Yf * (d.zdf - 1)) / (t2 - t4) / 1e9,
           up + 1, t2 - t4);
    dnsg_destroy(h);
    return 0;
}
