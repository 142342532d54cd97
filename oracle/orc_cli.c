/*
 * orc_cli.c -- command-line front end of the oracle (TEST INFRASTRUCTURE).
 *
 *   orc_cli init  <dir> key=value...   write the initial state as '-p' restart files
 *                                      (lbf-c/slf/suf/lin/uin/force .TSTR) into dir
 *   orc_cli run   <dir> key=value...   read the TSTR restart files from dir and run the
 *                                      reference driver loop (main:1254-1375), writing
 *                                      the same logs, turb-fld and lbf-* files
 *   orc_cli masks <file> key=value...  dump the reference-layout node codes (int32)
 *   orc_cli budget <dir> key=value...  the TKE-budget tool's per-snapshot pass
 *                                      (budget-no-slip-mgrd.c:82-92): for TSTR..TEND
 *                                      read <dir>/../vel-*, den-*.NNNN and write
 *                                      turb-fld-2d, vort-fluc-2d, KE-Budget-2d
 *                                      .000.NNNN (and ufl4/vfl4 every 200) into dir
 *   orc_cli uniform <file> key=value... write a synthetic uniform-grid lbf file
 *                                      (xdu, ydu, zdu, seed, u_max, noise) for '-i'
 *
 * init with init_kind=2 is the reference's '-i' start: it reads the uniform-grid
 * file lbf.TSTR from dir (sizes xdu, ydu, zdu) and interpolates it onto the grids.
 *
 * Keys are the dnsg_params field names (Rebs, ub, xdv, ydv, zdv, gratio, zds, CFL,
 * TSTR, TEND, ITER, pattern, gyplussize, ... , seed, u_max, noise, init_kind,
 * force_mode, g_const, tau_c, threads).
 */
#define _POSIX_C_SOURCE 200809L
#include "budget_oracle.h"
#include "dnsg_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int uni_size[3] = {0, 0, 0}; /* xdu, ydu, zdu (definitions.h:107-109) */

static int set_param(dnsg_params *p, const char *kv, int *threads) {
    char key[64];
    const char *eq = strchr(kv, '=');
    const char *val;
    size_t n;
    if (!eq) return -1;
    n = (size_t)(eq - kv);
    if (n >= sizeof key) return -1;
    memcpy(key, kv, n);
    key[n] = 0;
    val = eq + 1;
#define D(name) if (!strcmp(key, #name)) { p->name = strtod(val, NULL); return 0; }
#define I(name) if (!strcmp(key, #name)) { p->name = (int)strtol(val, NULL, 10); return 0; }
#define L(name) if (!strcmp(key, #name)) { p->name = strtol(val, NULL, 10); return 0; }
    D(Rebs) D(ub) I(zdv) I(ydv) I(xdv) I(TSTR) I(TEND) D(CFL) I(gratio) I(zds)
    I(gyplussize) I(wyplussize) I(gxplussize) I(wxplussize) I(pattern) L(ITER)
    I(force_mode) D(g_const) D(tau_c) I(precision) I(init_kind) D(u_max) D(noise) I(fltr)
#undef D
#undef I
#undef L
    if (!strcmp(key, "seed")) { p->seed = strtoull(val, NULL, 10); return 0; }
    if (!strcmp(key, "threads")) { *threads = (int)strtol(val, NULL, 10); return 0; }
    if (!strcmp(key, "xdu")) { uni_size[0] = (int)strtol(val, NULL, 10); return 0; }
    if (!strcmp(key, "ydu")) { uni_size[1] = (int)strtol(val, NULL, 10); return 0; }
    if (!strcmp(key, "zdu")) { uni_size[2] = (int)strtol(val, NULL, 10); return 0; }
    return -1;
}

static void defaults(dnsg_params *p) {
    memset(p, 0, sizeof *p);
    p->Rebs = 7200.; p->ub = 0.0866034; p->zdv = 32; p->ydv = 32; p->xdv = 64;
    p->TSTR = 0; p->TEND = 1; p->CFL = 0.1; p->gratio = 2; p->zds = 8;
    p->gyplussize = 56; p->wyplussize = 8; p->gxplussize = 4; p->wxplussize = 4;
    p->pattern = DNSG_PATTERN_NONE; p->ITER = 0; p->force_mode = DNSG_FORCE_FLUX;
    p->init_kind = DNSG_INIT_LAMINAR_NOISE; p->seed = 1234; p->u_max = 0.05; p->noise = 0.05;
}

static int read_all(const char *fn, double *dst, size_t n) {
    FILE *f = fopen(fn, "rb");
    int bad = !f || fread(dst, 8, n, f) != n;
    if (f) fclose(f);
    if (bad) fprintf(stderr, "cannot read %s\n", fn);
    return bad;
}

static int write_all(const char *fn, const double *src, size_t n) {
    FILE *f = fopen(fn, "wb");
    int bad = !f || fwrite(src, 8, n, f) != n;
    if (f) fclose(f);
    return bad;
}

/* the budget tool's snapshot loop (budget-no-slip-mgrd.c:82-92) */
static int budget(const dnsg_params *p, const char *dir) {
    const int GR = p->gratio, zdf = p->zds + 1, NZ = orc_budget_profile_len(p);
    const size_t nc = (size_t)p->xdv * p->ydv * p->zdv, nf = (size_t)GR * p->xdv * GR * p->ydv * zdf;
    double *vc = malloc(nc * 24), *dc = malloc(nc * 8), *vl = malloc(nf * 24), *dl = malloc(nf * 8);
    double *vu = malloc(nf * 24), *du = malloc(nf * 8), *turb = malloc((size_t)19 * NZ * 8);
    double *vort = malloc((size_t)3 * NZ * 8), *ke = malloc((size_t)30 * NZ * 8);
    const size_t npl = (size_t)GR * p->xdv * GR * p->ydv;
    double *ufl = malloc(npl * 8), *vfl = malloc(npl * 8);
    char fn[4096];
    int ts, rc = 0;
    if (!vc || !dc || !vl || !dl || !vu || !du || !turb || !vort || !ke || !ufl || !vfl) rc = 2;
    for (ts = p->TSTR; !rc && ts <= p->TEND; ts++) {
        const int dump = !(ts % 200);
#define RD(name, buf, n) (snprintf(fn, sizeof fn, "%s/../" name ".%.4d", dir, ts), read_all(fn, buf, n))
        rc = RD("vel-c", vc, 3 * nc) || RD("vel-slf", vl, 3 * nf) || RD("vel-suf", vu, 3 * nf) ||
             RD("den-c", dc, nc) || RD("den-slf", dl, nf) || RD("den-suf", du, nf);
#undef RD
        if (rc) break;
        if (orc_budget2d(p, vc, dc, vl, dl, vu, du, turb, vort, ke, dump ? ufl : NULL, dump ? vfl : NULL)) {
            rc = 2;
            break;
        }
        snprintf(fn, sizeof fn, "%s/turb-fld-2d.%.3d.%.4d", dir, 0, ts);
        rc |= write_all(fn, turb, (size_t)19 * NZ);
        snprintf(fn, sizeof fn, "%s/vort-fluc-2d.%.3d.%.4d", dir, 0, ts);
        rc |= write_all(fn, vort, (size_t)3 * NZ);
        snprintf(fn, sizeof fn, "%s/KE-Budget-2d.%.3d.%.4d", dir, 0, ts);
        rc |= write_all(fn, ke, (size_t)30 * NZ);
        if (dump) {
            snprintf(fn, sizeof fn, "%s/ufl4.%.4d", dir, ts);
            rc |= write_all(fn, ufl, npl);
            snprintf(fn, sizeof fn, "%s/vfl4.%.4d", dir, ts);
            rc |= write_all(fn, vfl, npl);
        }
    }
    free(vc); free(dc); free(vl); free(dl); free(vu); free(du); free(turb); free(vort); free(ke);
    free(ufl); free(vfl);
    return rc;
}

/* the '-i' start from dir/lbf.TSTR (main:982-997) */
static int init_uniform(orc_t *o, const char *dir, int tstr) {
    const size_t n = 2 * (size_t)uni_size[0] * uni_size[1] * (uni_size[2] + 2) * 19;
    char fn[4096], err[256];
    double *u;
    FILE *f;
    int rc;
    snprintf(fn, sizeof fn, "%s/lbf.%.4d", dir, tstr); /* initializer.c:303 */
    if (!n || !(u = malloc(n * 8))) return 2;
    f = fopen(fn, "rb");
    rc = !f || fread(u, 8, n, f) != n;
    if (f) fclose(f);
    if (!rc && orc_init_uniform(o, u, uni_size[0], uni_size[1], uni_size[2], err, sizeof err)) {
        fprintf(stderr, "%s\n", err);
        rc = 1;
    }
    free(u);
    return rc;
}

int main(int argc, char **argv) {
    dnsg_params p;
    orc_t *o;
    char err[256] = "";
    int i, threads = 0, rc = 0;
    if (argc < 3) {
        fprintf(stderr, "usage: %s init|run|masks <dir|file> key=value...\n", argv[0]);
        return 2;
    }
    defaults(&p);
    for (i = 3; i < argc; i++)
        if (set_param(&p, argv[i], &threads)) {
            fprintf(stderr, "bad parameter '%s'\n", argv[i]);
            return 2;
        }
    if (threads > 0) orc_set_threads(threads);
    if (!strcmp(argv[1], "uniform")) {
        const size_t n = 2 * (size_t)uni_size[0] * uni_size[1] * (uni_size[2] + 2) * 19;
        double *u = n ? malloc(n * 8) : NULL;
        FILE *f;
        if (!u) return 2;
        orc_uniform_field(uni_size[0], uni_size[1], uni_size[2], p.seed, p.u_max, p.noise, u);
        f = fopen(argv[2], "wb");
        rc = !f || fwrite(u, 8, n, f) != n;
        if (f) fclose(f);
        free(u);
        return rc ? 1 : 0;
    }
    if (!strcmp(argv[1], "budget")) return budget(&p, argv[2]) ? 1 : 0;
    o = orc_create(&p, err, sizeof err);
    if (!o) {
        fprintf(stderr, "orc_create: %s\n", err);
        return 2;
    }
    if (!strcmp(argv[1], "init")) {
        if (p.init_kind == DNSG_INIT_REST) orc_init_rest(o);
        else if (p.init_kind == 2) rc = init_uniform(o, argv[2], p.TSTR);
        else orc_init_seeded(o);
        if (p.force_mode == DNSG_FORCE_FLUX) orc_set_force(o, 0.);
        if (!rc) rc = orc_write_checkpoint(o, argv[2], p.TSTR);
    } else if (!strcmp(argv[1], "run")) {
        dnsg_dims d;
        orc_get_dims(o, &d);
        rc = orc_read_checkpoint(o, argv[2], p.TSTR);
        if (!rc) {
            printf("Average Density in the channel at start: %.12f\n", orc_mass(o));
            rc = orc_run(o, argv[2], d.up, 1);
            printf("Average Density in the channel at end: %.12f\n", orc_mass(o));
        }
    } else if (!strcmp(argv[1], "masks")) {
        dnsg_dims d;
        int *sc, *sl, *su;
        size_t nc, nf;
        FILE *f;
        orc_get_dims(o, &d);
        nc = (size_t)(d.Xc + 1) * (d.Yc + 1) * d.zdc;
        nf = (size_t)(d.Xf + 1) * (d.Yf + 1) * d.zdf;
        sc = malloc(nc * sizeof(int)); sl = malloc(nf * sizeof(int)); su = malloc(nf * sizeof(int));
        orc_ref_masks(o, sc, sl, su);
        f = fopen(argv[2], "wb");
        if (!f) rc = DNSG_ERR_IO;
        else {
            fwrite(sc, sizeof(int), nc, f);
            fwrite(sl, sizeof(int), nf, f);
            fwrite(su, sizeof(int), nf, f);
            fclose(f);
        }
        free(sc); free(sl); free(su);
    } else {
        fprintf(stderr, "unknown command %s\n", argv[1]);
        rc = 2;
    }
    orc_destroy(o);
    return rc ? 1 : 0;
}
