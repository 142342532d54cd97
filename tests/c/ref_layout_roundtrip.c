/*
 * ref_layout_roundtrip.c -- the reference-side binding of include/dnsg.h, exercised
 * with arrays allocated exactly as the reference's main allocates them
 * (DNS-solver-with-grid-embedding/src/dns-q-mgrd-all-patterns-xyc.c:145-151,
 * 473-883): V.ftemp [xdc][ydc][zdc][19] with xdc = 1+xdv, V.slftemp / V.suftemp
 * [xdf][ydf][zdf][19] with xdf = GR(xdc-1)+1, and the six interface levels
 * 23(ydc+5)(xdc+5) doubles each, addressed with the Fbi macro (definitions.h:132).
 *
 * 1. Fill the interiors with a seeded state (equilibria of a perturbed channel
 *    profile), upload it with dnsg_upload_state_ghosted, download it with
 *    dnsg_download_state_ghosted into fresh arrays: every interior value comes back
 *    bit for bit and every ghost cell holds its periodic image.
 * 2. Upload the same state through the file-layout entry point (dnsg_upload_state)
 *    into a second handle; step both handles 4 coarse steps; the two states, read
 *    back through either entry point, agree bit for bit.
 * Exit status 0 on success.  Needs a GPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dnsg.h"

/* reference macros (definitions.h:130-132), here as functions */
static size_t Fb(int i, int j, int k, int b, int yd, int zd) { return (((size_t)i * yd + j) * zd + k) * 19 + b; }
static size_t Fbi(int i, int j, int b, int yd) { return ((size_t)(i + 2) * (yd + 5) + (j + 2)) * 23 + b; }

static const double W[19] = {1. / 3, 1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 18, 1. / 36, 1. / 36, 1. / 36,
                             1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36, 1. / 36};
static const int E0[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0};
static const int E1[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
static const int E2[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

static int fails = 0;
#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            fprintf(stderr, "FAIL: " __VA_ARGS__); \
            fputc('\n', stderr);               \
            if (++fails > 20) exit(1);         \
        }                                      \
    } while (0)

static double noise(unsigned long n) {  /* deterministic value in [-1, 1) */
    n = (n ^ (n >> 17)) * 0x9E3779B97F4A7C15UL;
    return (double)((n >> 11) & ((1UL << 40) - 1)) / (double)(1UL << 39) - 1.;
}

static void feq(double rho, double u, double v, double w, double *f) {
    int q;
    for (q = 0; q < 19; q++) {
        double eu = E0[q] * u + E1[q] * v + E2[q] * w;
        f[q] = W[q] * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * (u * u + v * v + w * w));
    }
}

static void fill_lattice(double *a, int X, int Y, int Z, unsigned long tag) {
    int i, j, k;
    for (i = 0; i < X; i++)
        for (j = 0; j < Y; j++)
            for (k = 0; k < Z; k++) {
                unsigned long n = (((tag * 131 + i) * 131 + j) * 131 + k) * 4;
                double z = (k + 0.5) / Z;
                feq(1. + 0.01 * noise(n), 0.04 * z * (1. - z) + 0.002 * noise(n + 1), 0.002 * noise(n + 2),
                    0.002 * noise(n + 3), &a[Fb(i, j, k, 0, Y + 1, Z)]);
            }
}

static void fill_level(double *a, int X, int Y, unsigned long tag) {
    int i, j, b;
    for (i = 0; i < X; i++)
        for (j = 0; j < Y; j++)
            for (b = 0; b < 23; b++) a[Fbi(i, j, b, Y + 1)] = (b == 0 ? 1. : 0.) + 1e-3 * noise((tag * 977 + i) * 977 * 977 + j * 977 + b);
}

/* interior bitwise and ghost plane/row = periodic image */
static void cmp_lattice(const double *a, const double *b, int X, int Y, int Z, const char *what) {
    int i, j, k, q;
    for (i = 0; i <= X; i++)
        for (j = 0; j <= Y; j++)
            for (k = 0; k < Z; k++)
                for (q = 0; q < 19; q++) {
                    double want = a[Fb(i % X, j % Y, k, q, Y + 1, Z)];
                    double got = b[Fb(i, j, k, q, Y + 1, Z)];
                    CHECK(memcmp(&want, &got, 8) == 0, "%s (%d,%d,%d,%d): %.17g != %.17g", what, i, j, k, q, got, want);
                }
}

static void cmp_level(const double *a, const double *b, int X, int Y, const char *what) {
    int i, j, v;
    for (i = -2; i < X + 4; i++)
        for (j = -2; j < Y + 4; j++)
            for (v = 0; v < 23; v++) {
                double want = a[Fbi((i + X) % X, (j + Y) % Y, v, Y + 1)];
                double got = b[Fbi(i, j, v, Y + 1)];
                CHECK(memcmp(&want, &got, 8) == 0, "%s (%d,%d,%d): %.17g != %.17g", what, i, j, v, got, want);
            }
}

static void die(dnsg_t *h, const char *what, int rc) {
    fprintf(stderr, "%s failed: %d %s\n", what, rc, h ? dnsg_last_error(h) : "");
    exit(1);
}

int main(void) {
    dnsg_params p;
    dnsg_dims d;
    dnsg_t *h1 = NULL, *h2 = NULL;
    int rc, l, xdc, ydc, zdc, xdf, ydf, zdf;
    size_t nc, nf, ni;
    double *ftemp, *slftemp, *suftemp, *lev[6], *back[9], *file[5], *file2[5], f1, f2;
    dnsg_params_default(&p);
    p.xdv = 16; p.ydv = 8; p.zdv = 12; p.gratio = 2; p.zds = 4;
    p.Rebs = 20.0; p.ub = 0.0333333333333; p.CFL = 5.0;
    p.pattern = DNSG_PATTERN_RSLIP; p.gyplussize = 4; p.wyplussize = 4;
    p.force_mode = DNSG_FORCE_CONST; p.g_const = 2e-5;
    if ((rc = dnsg_dims_from_params(&p, &d))) die(NULL, "dnsg_dims_from_params", rc);
    /* main:145-151 on one rank (XDIM = YDIM = 1) */
    xdc = 1 + p.xdv; ydc = 1 + p.ydv; zdc = p.zdv;
    xdf = p.gratio * (xdc - 1) + 1; ydf = p.gratio * (ydc - 1) + 1; zdf = p.zds + 1;
    nc = (size_t)xdc * ydc * zdc * 19;
    nf = (size_t)xdf * ydf * zdf * 19;
    ni = (size_t)23 * (ydc + 5) * (xdc + 5);
    /* main:473-883 */
    ftemp = calloc(nc, sizeof(double));
    slftemp = calloc(nf, sizeof(double));
    suftemp = calloc(nf, sizeof(double));
    for (l = 0; l < 6; l++) lev[l] = calloc(ni, sizeof(double));
    for (l = 0; l < 3; l++) back[l] = calloc(l == 0 ? nc : nf, sizeof(double));
    for (l = 3; l < 9; l++) back[l] = calloc(ni, sizeof(double));
    file[0] = malloc(d.n_coarse * 8); file[1] = malloc(d.n_band * 8); file[2] = malloc(d.n_band * 8);
    file[3] = malloc(d.n_plane * 8); file[4] = malloc(d.n_plane * 8);
    file2[0] = malloc(d.n_coarse * 8); file2[1] = malloc(d.n_band * 8); file2[2] = malloc(d.n_band * 8);
    file2[3] = malloc(d.n_plane * 8); file2[4] = malloc(d.n_plane * 8);
    fill_lattice(ftemp, xdc - 1, ydc - 1, zdc, 1);
    fill_lattice(slftemp, xdf - 1, ydf - 1, zdf, 2);
    fill_lattice(suftemp, xdf - 1, ydf - 1, zdf, 3);
    for (l = 0; l < 6; l++) fill_level(lev[l], xdc - 1, ydc - 1, 10 + l);

    if ((rc = dnsg_create(&p, &h1))) die(h1, "dnsg_create", rc);
    {
        const double *lin[3] = {lev[0], lev[1], lev[2]}, *uin[3] = {lev[3], lev[4], lev[5]};
        double *blin[3] = {back[3], back[4], back[5]}, *buin[3] = {back[6], back[7], back[8]};
        if ((rc = dnsg_upload_state_ghosted(h1, ftemp, slftemp, suftemp, lin, uin, 0.))) die(h1, "upload_ghosted", rc);
        if ((rc = dnsg_download_state_ghosted(h1, back[0], back[1], back[2], blin, buin, &f1)))
            die(h1, "download_ghosted", rc);
    }
    cmp_lattice(ftemp, back[0], xdc - 1, ydc - 1, zdc, "V.ftemp");
    cmp_lattice(slftemp, back[1], xdf - 1, ydf - 1, zdf, "V.slftemp");
    cmp_lattice(suftemp, back[2], xdf - 1, ydf - 1, zdf, "V.suftemp");
    for (l = 0; l < 6; l++) cmp_level(lev[l], back[3 + l], xdc - 1, ydc - 1, "interface level");
    CHECK(f1 == p.g_const, "force %.17g", f1);

    /* the same state through the file layout into a second handle */
    if ((rc = dnsg_download_state(h1, file[0], file[1], file[2], file[3], file[4], NULL))) die(h1, "download", rc);
    if ((rc = dnsg_create(&p, &h2))) die(h2, "dnsg_create", rc);
    if ((rc = dnsg_upload_state(h2, file[0], file[1], file[2], file[3], file[4], 0.))) die(h2, "upload", rc);
    if ((rc = dnsg_step(h1, 4, NULL))) die(h1, "dnsg_step", rc);
    if ((rc = dnsg_step(h2, 4, NULL))) die(h2, "dnsg_step", rc);
    if ((rc = dnsg_download_state(h1, file[0], file[1], file[2], file[3], file[4], &f1))) die(h1, "download", rc);
    if ((rc = dnsg_download_state(h2, file2[0], file2[1], file2[2], file2[3], file2[4], &f2))) die(h2, "download", rc);
    CHECK(memcmp(file[0], file2[0], d.n_coarse * 8) == 0, "coarse after 4 steps");
    CHECK(memcmp(file[1], file2[1], d.n_band * 8) == 0, "lower band after 4 steps");
    CHECK(memcmp(file[2], file2[2], d.n_band * 8) == 0, "upper band after 4 steps");
    CHECK(memcmp(file[3], file2[3], d.n_plane * 8) == 0, "lower levels after 4 steps");
    CHECK(memcmp(file[4], file2[4], d.n_plane * 8) == 0, "upper levels after 4 steps");
    CHECK(f1 == f2, "force after 4 steps");
    {   /* the ghosted download of the stepped state: interior = file layout */
        double *blin[3] = {back[3], back[4], back[5]}, *buin[3] = {back[6], back[7], back[8]};
        int i, j, k, q;
        if ((rc = dnsg_download_state_ghosted(h2, back[0], back[1], back[2], blin, buin, NULL)))
            die(h2, "download_ghosted", rc);
        for (i = 0; i < xdc - 1; i++)
            for (j = 0; j < ydc - 1; j++)
                for (k = 0; k < zdc; k++)
                    for (q = 0; q < 19; q++)
                        CHECK(back[0][Fb(i, j, k, q, ydc, zdc)] == file[0][(((size_t)i * (ydc - 1) + j) * zdc + k) * 19 + q],
                              "stepped coarse (%d,%d,%d,%d)", i, j, k, q);
        for (l = 0; l < 3; l++)
            for (i = 0; i < xdc - 1; i++)
                for (j = 0; j < ydc - 1; j++)
                    for (q = 0; q < 23; q++)
                        CHECK(back[3 + l][Fbi(i, j, q, ydc)] ==
                                  file[3][(size_t)l * d.Xc * d.Yc * 23 + ((size_t)i * d.Yc + j) * 23 + q],
                              "stepped level %d (%d,%d,%d)", l, i, j, q);
    }
    dnsg_destroy(h1);
    dnsg_destroy(h2);
    if (fails) {
        fprintf(stderr, "%d failures\n", fails);
        return 1;
    }
    printf("reference layouts: round trip and 4 steps bit for bit\n");
    return 0;
}
