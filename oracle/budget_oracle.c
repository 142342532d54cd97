/*
 * budget_oracle.c -- CPU restatement of the reference's per-snapshot turbulence
 * statistics and kinetic-energy budget (TEST INFRASTRUCTURE: the checker of the
 * HIP budget, never the product path).
 *
 * Follows TKE-budget/computations: reading.c (the snapshot arrays with two
 * periodic ghost planes on each side in x), budget2D.c:13-1736 (the statistics)
 * and uv_fluc.c:14-46 (the fluctuation plane dump), on one rank.  The arithmetic
 * keeps the reference's order (sums in its loop order, every expression grouped
 * as written), so the outputs are byte-comparable with the reference tool's files.
 *
 * Profiles are NZ = 2 zdf + zdc - 4 long: lower band k = 0 .. zdf-1, coarse
 * k = 2 .. zdc-3 at zdf-2+k, upper band k = 0 .. zdf-1 at zdf+zdc-4+k.
 */
#include "budget_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* element (i, j, k, b) of a snapshot array with two ghost planes before x = 0 and d
 * values per node: the reference's Fb with its +2 plane offset (definitions.h of
 * the tool).  k outside 0 .. zd-1 reaches the neighbouring row, as it does there. */
static inline double at(const double *a, int i, int j, int k, int b, int yd, int zd, int d) {
    return a[(((long)i + 2) * yd * zd + (long)j * zd + k) * d + b];
}

/* snapshot [X][Y][Z][d] -> [X+4][Y][Z][d] with periodic ghost planes (reading.c:33-57) */
static double *pad_x(const double *src, int X, int Y, int Z, int d) {
    const size_t plane = (size_t)Y * Z * d;
    double *p = malloc(plane * (X + 4) * sizeof(double));
    int i;
    if (!p) return NULL;
    for (i = -2; i < X + 2; i++) memcpy(p + (size_t)(i + 2) * plane, src + (size_t)((i + X) % X) * plane, plane * 8);
    return p;
}

enum { /* the per-node accumulators, in the order budget2D.c:449-617 updates them */
    A_UU, A_UV, A_UW, A_VV, A_VW, A_WW, A_SKU, A_SKV, A_SKW, A_SKP, A_SKUW, A_SKUV, A_SKVW,
    A_KUU, A_KUV, A_KUW, A_KUP, A_KUUW, A_KUVW, A_KUUV, A_DD, A_UUW, A_VVW, A_WWW, A_UWW,
    A_PU, A_PV, A_PW, A_WX, A_WY, A_WZ, A_PAIUU, A_PAIVV, A_PAIWW, A_PAIUW,
    A_USUW, A_VSVW, A_WSWW, A_WSUW, A_USWW, A_EPUU, A_EPVV, A_EPWW, A_EPUW, NACC
};

typedef struct {
    int Xc, Yc, zdc, Xf, Yf, zdf, GR, NZ;
    double coef, nu;
    double *vc, *vl, *vu;       /* padded fluctuating velocities */
    const double *dc, *dl, *du; /* padded densities */
} bud_t;

/* z derivative of component b at (i, j, k) of grid g (0 coarse, 1 lower, 2 upper),
   budget2D.c:539-593 (lower), :720-748 (coarse), :904-940 (upper) */
static double ddz(const bud_t *B, int g, int i, int j, int k, int b) {
    const int GR = B->GR;
    const double coef = B->coef;
    if (g == 1) {
        const double *v = B->vl;
        const int yd = B->Yf, zd = B->zdf;
#define V(kk) at(v, i, j, (kk), b, yd, zd, 3)
        if (k == 1) {
            const double tmp = (2. / 3.) * (V(k + 3) - V(k + 1)) - (1. / 12.) * (V(k + 4) - V(k));
            return -(7. / 3.) * V(k) + 6. * V(k + 1) - 3. * V(k + 2) - (2. / 3.) * V(k + 3) + 3. * tmp;
        } else if (k == zd - 1) {
            const double c = at(B->vc, i / GR, j / GR, 2, b, B->Yc, B->zdc, 3);
            return (coef * (c - V(k)) + GR * (V(k) - V(k - 1))) * 1. / (1. + GR);
        }
        return 0.5 * (V(k + 1) - V(k - 1));
#undef V
    } else if (g == 0) {
        const double *v = B->vc;
        const int yd = B->Yc, zd = B->zdc;
#define V(kk) at(v, i, j, (kk), b, yd, zd, 3)
        if (k == 2) return 0.5 * coef * (V(k + 1) - at(B->vl, i * GR, j * GR, B->zdf - 1, b, B->Yf, B->zdf, 3));
        if (k == 3 || k == B->zdc - 4) return 0.5 * coef * (V(k + 1) - V(k - 1));
        if (k == B->zdc - 3) return 0.5 * coef * (at(B->vu, i * GR, j * GR, 0, b, B->Yf, B->zdf, 3) - V(k - 1));
        return 0.5 * coef * (V(k + 1) - V(k - 1));
#undef V
    } else {
        const double *v = B->vu;
        const int yd = B->Yf, zd = B->zdf;
#define V(kk) at(v, i, j, (kk), b, yd, zd, 3)
        if (k == 0) {
            const double c = at(B->vc, i / GR, j / GR, B->zdc - 3, b, B->Yc, B->zdc, 3);
            return (GR * (V(k + 1) - V(k)) + coef * (V(k) - c)) * 1. / (1. + GR);
        } else if (k == zd - 2) {
            const double tmp = (2. / 3.) * (V(k - 1) - V(k - 3)) - (1. / 12.) * (V(k) - V(k - 4));
            return (7. / 3.) * V(k) - 6. * V(k - 1) + 3. * V(k - 2) + (2. / 3.) * V(k - 3) + 3. * tmp;
        }
        return 0.5 * (V(k + 1) - V(k - 1));
#undef V
    }
}

/* one grid's per-node pass (budget2D.c:449-617, :636-797, :800-987) into acc[NACC][NZ] */
static void node_pass(const bud_t *B, int g, const double *davg, double *acc) {
    const int X = g == 0 ? B->Xc : B->Xf, Y = g == 0 ? B->Yc : B->Yf, Z = g == 0 ? B->zdc : B->zdf;
    const int k0 = g == 1 ? 1 : (g == 0 ? 2 : 0), k1 = g == 1 ? Z : (g == 0 ? Z - 2 : Z - 1);
    const int base = g == 1 ? 0 : (g == 0 ? B->zdf - 2 : B->zdf + B->zdc - 4);
    const double *v = g == 0 ? B->vc : (g == 1 ? B->vl : B->vu), *dn = g == 0 ? B->dc : (g == 1 ? B->dl : B->du);
    const double cx = g == 0 ? B->coef : 1.; /* the coarse x/y derivatives carry coef */
    int i, j, k;
    for (i = 0; i < X; i++)
        for (j = 0; j < Y; j++)
            for (k = k0; k < k1; k++) {
                double *A = acc + base + k;
                const int NZ = B->NZ;
                const double u = at(v, i, j, k, 0, Y, Z, 3), vv = at(v, i, j, k, 1, Y, Z, 3),
                             w = at(v, i, j, k, 2, Y, Z, 3);
                const double dens = at(dn, i, j, k, 0, Y, Z, 1);
                const double dfl = (1. / 3.) * dens - davg[base + k];
                /* j neighbours (:495-525): periodic */
                const int jp = j == 0 ? Y - 1 : j - 1, jpp = j == 0 ? Y - 2 : (j == 1 ? Y - 1 : j - 2);
                const int jn = j == Y - 1 ? 0 : j + 1, jnn = j == Y - 1 ? 1 : (j == Y - 2 ? 0 : j + 2);
                double d[3][3]; /* d[c][axis]: d(u_c)/d(x, y, z) */
                int c;
                for (c = 0; c < 3; c++) {
                    const double gx = (2. / 3.) * (at(v, i + 1, j, k, c, Y, Z, 3) - at(v, i - 1, j, k, c, Y, Z, 3)) -
                                      (1. / 12.) * (at(v, i + 2, j, k, c, Y, Z, 3) - at(v, i - 2, j, k, c, Y, Z, 3));
                    const double gy = (2. / 3.) * (at(v, i, jn, k, c, Y, Z, 3) - at(v, i, jp, k, c, Y, Z, 3)) -
                                      (1. / 12.) * (at(v, i, jnn, k, c, Y, Z, 3) - at(v, i, jpp, k, c, Y, Z, 3));
                    d[c][0] = g == 0 ? cx * gx : gx;
                    d[c][1] = g == 0 ? cx * gy : gy;
                    d[c][2] = ddz(B, g, i, j, k, c);
                }
                {
                    const double dudx = d[0][0], dudy = d[0][1], dudz = d[0][2];
                    const double dvdx = d[1][0], dvdy = d[1][1], dvdz = d[1][2];
                    const double dwdx = d[2][0], dwdy = d[2][1], dwdz = d[2][2];
                    const double ox = (dwdy - dvdz), oy = (dudz - dwdx), oz = (dvdx - dudy);
#define ACC(n) A[(size_t)(n) * NZ]
                    ACC(A_UU) += u * u; ACC(A_UV) += u * vv; ACC(A_UW) += u * w;
                    ACC(A_VV) += vv * vv; ACC(A_VW) += vv * w; ACC(A_WW) += w * w;
                    ACC(A_SKU) += u * u * u; ACC(A_SKV) += vv * vv * vv; ACC(A_SKW) += w * w * w;
                    ACC(A_SKP) += dfl * dfl * dfl;
                    ACC(A_SKUW) += u * u * u * w * w * w; ACC(A_SKUV) += u * u * u * vv * vv * vv;
                    ACC(A_SKVW) += vv * vv * vv * w * w * w;
                    ACC(A_KUU) += u * u * u * u; ACC(A_KUV) += vv * vv * vv * vv; ACC(A_KUW) += w * w * w * w;
                    ACC(A_KUP) += dfl * dfl * dfl * dfl;
                    ACC(A_KUUW) += u * u * u * u * w * w * w * w; ACC(A_KUVW) += vv * vv * vv * vv * w * w * w * w;
                    ACC(A_KUUV) += u * u * u * u * vv * vv * vv * vv;
                    ACC(A_DD) += dfl * dfl;
                    ACC(A_UUW) += u * u * w; ACC(A_VVW) += vv * vv * w; ACC(A_WWW) += w * w * w; ACC(A_UWW) += u * w * w;
                    ACC(A_PU) += u * dfl; ACC(A_PV) += vv * dfl; ACC(A_PW) += w * dfl;
                    ACC(A_WX) += ox * ox; ACC(A_WY) += oy * oy; ACC(A_WZ) += oz * oz;
                    ACC(A_PAIUU) += dfl * dudx; ACC(A_PAIVV) += dfl * dvdy; ACC(A_PAIWW) += dfl * dwdz;
                    ACC(A_PAIUW) += dfl * (dudz + dwdx);
                    ACC(A_USUW) += 0.5 * u * (dudz + dwdx); ACC(A_VSVW) += 0.5 * vv * (dvdz + dwdy);
                    ACC(A_WSWW) += 0.5 * w * (dwdz + dwdz); ACC(A_WSUW) += 0.5 * w * (dudz + dwdx);
                    ACC(A_USWW) += 0.5 * u * (dwdz + dwdz);
                    ACC(A_EPUU) += dudx * (dudx + dudx) + dudy * (dudy + dvdx) + dudz * (dudz + dwdx);
                    ACC(A_EPVV) += dvdx * (dvdx + dudy) + dvdy * (dvdy + dvdy) + dvdz * (dvdz + dwdy);
                    ACC(A_EPWW) += dwdx * (dwdx + dudz) + dwdy * (dwdy + dvdz) + dwdz * (dwdz + dwdz);
                    ACC(A_EPUW) += dudx * (dwdx + dudz) + dudy * (dwdy + dvdz) + dudz * (dwdz + dwdz) +
                                   dwdx * (dudx + dudx) + dwdy * (dudy + dvdx) + dwdz * (dudz + dwdx);
#undef ACC
                }
            }
}

/* mean of one grid's layers (budget2D.c:177-244): sums in j, k, i order */
static void layer_means(const double *v, const double *dn, int X, int Y, int Z, int k0, int k1, int base,
                        double *U, double *Vm, double *W, double *D) {
    int i, j, k;
    for (j = 0; j < Y; j++)
        for (k = k0; k < k1; k++)
            for (i = 0; i < X; i++) {
                U[base + k] += at(v, i, j, k, 0, Y, Z, 3);
                Vm[base + k] += at(v, i, j, k, 1, Y, Z, 3);
                W[base + k] += at(v, i, j, k, 2, Y, Z, 3);
                D[base + k] += (1. / 3.) * at(dn, i, j, k, 0, Y, Z, 1);
            }
    for (k = k0; k < k1; k++) {
        U[base + k] /= (double)(X * Y);
        Vm[base + k] /= (double)(X * Y);
        W[base + k] /= (double)(X * Y);
        D[base + k] /= (double)(X * Y);
    }
}

static void subtract_means(double *v, int X, int Y, int Z, int k0, int k1, int base, const double *U,
                           const double *Vm, const double *W) {
    int i, j, k;
    for (i = -2; i < X + 2; i++) /* :265-309, ghost planes included */
        for (j = 0; j < Y; j++)
            for (k = k0; k < k1; k++) {
                double *p = v + (((long)i + 2) * Y * Z + (long)j * Z + k) * 3;
                p[0] -= U[base + k];
                p[1] -= Vm[base + k];
                p[2] -= W[base + k];
            }
}

int orc_budget_profile_len(const dnsg_params *p) { return 2 * (p->zds + 1) + p->zdv - 4; }

int orc_budget2d(const dnsg_params *p, const double *velc, const double *denc, const double *velfl,
                 const double *denfl, const double *velfu, const double *denfu, double *turb, double *vort,
                 double *ke, double *ufl, double *vfl) {
    bud_t B;
    const int GR = p->gratio, zdf = p->zds + 1, zdc = p->zdv;
    const int NZ = 2 * zdf + (zdc - 4), NZs = 2 * zdf + GR * (zdc - 4) + GR - 1;
    double *U = calloc((size_t)NZ * 4, 8), *Vm, *W, *D, *dUdz, *acc, *pr;
    const double tauf = (3. / p->Rebs) * p->ub * (NZs - 2.) + 0.5; /* budget2D.c:55-58 */
    double *dc = NULL, *dl = NULL, *du = NULL;
    int k, rc = 0;
    memset(&B, 0, sizeof B);
    B.Xc = p->xdv; B.Yc = p->ydv; B.zdc = zdc; B.Xf = GR * p->xdv; B.Yf = GR * p->ydv; B.zdf = zdf;
    B.GR = GR; B.NZ = NZ;
    B.nu = (1. / 6.) * (2. * tauf - 1.);
    B.coef = (1. / ((double)GR));
    acc = calloc((size_t)NACC * NZ, 8);
    dUdz = calloc(NZ, 8);
    pr = calloc((size_t)NZ * 40, 8); /* derived profiles */
    B.vc = pad_x(velc, B.Xc, B.Yc, zdc, 3); B.vl = pad_x(velfl, B.Xf, B.Yf, zdf, 3); B.vu = pad_x(velfu, B.Xf, B.Yf, zdf, 3);
    dc = pad_x(denc, B.Xc, B.Yc, zdc, 1); dl = pad_x(denfl, B.Xf, B.Yf, zdf, 1); du = pad_x(denfu, B.Xf, B.Yf, zdf, 1);
    B.dc = dc; B.dl = dl; B.du = du;
    if (!U || !acc || !dUdz || !pr || !B.vc || !B.vl || !B.vu || !dc || !dl || !du) { rc = -1; goto out; }
    Vm = U + NZ; W = U + 2 * NZ; D = U + 3 * NZ;
    layer_means(B.vl, dl, B.Xf, B.Yf, zdf, 0, zdf, 0, U, Vm, W, D);
    layer_means(B.vc, dc, B.Xc, B.Yc, zdc, 2, zdc - 2, zdf - 2, U, Vm, W, D);
    layer_means(B.vu, du, B.Xf, B.Yf, zdf, 0, zdf, zdf + zdc - 4, U, Vm, W, D);
    subtract_means(B.vl, B.Xf, B.Yf, zdf, 0, zdf, 0, U, Vm, W);
    subtract_means(B.vc, B.Xc, B.Yc, zdc, 2, zdc - 2, zdf - 2, U, Vm, W);
    subtract_means(B.vu, B.Xf, B.Yf, zdf, 0, zdf, zdf + zdc - 4, U, Vm, W);
    /* dU/dz of the mean profile (:311-444) */
    {
        const double coef = B.coef;
        int zd = zdf;
        for (k = 1; k < zd; k++) {
            const double *a = U;
            if (k == 1) {
                const double tmp = (2. / 3.) * (a[k + 3] - a[k + 1]) - (1. / 12.) * (a[k + 4] - a[k]);
                dUdz[k] = -(7. / 3.) * a[k] + 6. * a[k + 1] - 3. * a[k + 2] - (2. / 3.) * a[k + 3] + 3. * tmp;
            } else if (k == zd - 1) dUdz[k] = (coef * (a[k + 1] - a[k]) + GR * (a[k] - a[k - 1])) * 1. / (1. + GR);
            else if (k == 2 || k == zd - 2) dUdz[k] = 0.5 * (a[k + 1] - a[k - 1]);
            else dUdz[k] = (2. / 3.) * (a[k + 1] - a[k - 1]) - (1. / 12.) * (a[k + 2] - a[k - 2]);
        }
        for (k = 2; k < zdc - 2; k++) {
            const double *a = U + zdf - 2;
            if (k == 2 || k == 3 || k == zdc - 4 || k == zdc - 3) dUdz[zdf - 2 + k] = 0.5 * coef * (a[k + 1] - a[k - 1]);
            else dUdz[zdf - 2 + k] = coef * ((2. / 3.) * (a[k + 1] - a[k - 1]) - (1. / 12.) * (a[k + 2] - a[k - 2]));
        }
        for (k = 0; k < zd - 1; k++) {
            const double *a = U + zdf + zdc - 4;
            double *o = dUdz + zdf + zdc - 4;
            if (k == 0) o[k] = (GR * (a[k + 1] - a[k]) + coef * (a[k] - a[k - 1])) * 1. / (1. + GR);
            else if (k == zd - 2) {
                const double tmp = (2. / 3.) * (a[k - 1] - a[k - 3]) - (1. / 12.) * (a[k] - a[k - 4]);
                o[k] = (7. / 3.) * a[k] - 6. * a[k - 1] + 3. * a[k - 2] + (2. / 3.) * a[k - 3] + 3. * tmp;
            } else if (k == 1 || k == zd - 3) o[k] = 0.5 * (a[k + 1] - a[k - 1]);
            else o[k] = (2. / 3.) * (a[k + 1] - a[k - 1]) - (1. / 12.) * (a[k + 2] - a[k - 2]);
        }
    }
    node_pass(&B, 1, D, acc);
    node_pass(&B, 0, D, acc);
    node_pass(&B, 2, D, acc);
    orc_budget_finish(p, U, dUdz, acc, turb, vort, ke);
    if (ufl && vfl) { /* uv_fluc.c:25-33: the lower band's fluctuations at k = 4 */
        int i, j;
        for (i = 0; i < B.Xf; i++)
            for (j = 0; j < B.Yf; j++) {
                ufl[(size_t)i * B.Yf + j] = at(B.vl, i, j, 4, 0, B.Yf, zdf, 3);
                vfl[(size_t)i * B.Yf + j] = at(B.vl, i, j, 4, 1, B.Yf, zdf, 3);
            }
    }
out:
    free(U); free(acc); free(dUdz); free(pr);
    free(B.vc); free(B.vl); free(B.vu); free(dc); free(dl); free(du);
    return rc;
}

/* normalisation and the derived profiles (budget2D.c:1003-1547) from the means
 * U[4][NZ] (U, V, W, rho/3), dU/dz and the raw sums acc[NACC][NZ] */
void orc_budget_finish(const dnsg_params *p, const double *Um, const double *dUdz, const double *acc_in,
                       double *turb, double *vort, double *ke) {
    const int GR = p->gratio, zdf = p->zds + 1, zdc = p->zdv;
    const int NZ = 2 * zdf + (zdc - 4), NZs = 2 * zdf + GR * (zdc - 4) + GR - 1;
    const double tauf = (3. / p->Rebs) * p->ub * (NZs - 2.) + 0.5;
    const double nu = (1. / 6.) * (2. * tauf - 1.), coef = (1. / ((double)GR));
    const double nf = (double)(GR * p->xdv) * (GR * p->ydv), nc = (double)p->xdv * p->ydv;
    double *a = malloc((size_t)NACC * NZ * 8);
    double *UTIN = turb + 3 * NZ, *VTIN = turb + 4 * NZ, *WTIN = turb + 5 * NZ, *DTIN = turb + 8 * NZ;
    double *KE = ke + 29 * NZ;
    double *Pii = ke, *Puu = ke + NZ, *Pvv = ke + 2 * NZ, *Pww = ke + 3 * NZ, *Puw = ke + 4 * NZ;
    double *Tiir = ke + 5 * NZ, *Tuur = ke + 6 * NZ, *Tvvr = ke + 7 * NZ, *Twwr = ke + 8 * NZ, *Tuwr = ke + 9 * NZ;
    double *Tiipai = ke + 10 * NZ, *Tuupai = ke + 11 * NZ, *Tvvpai = ke + 12 * NZ, *Twwpai = ke + 13 * NZ,
           *Tuwpai = ke + 14 * NZ;
    double *Tiis = ke + 19 * NZ, *Tuus = ke + 20 * NZ, *Tvvs = ke + 21 * NZ, *Twws = ke + 22 * NZ, *Tuws = ke + 23 * NZ;
    double *Epii = ke + 24 * NZ;
    int k, n;
    if (!a) return;
    memcpy(a, acc_in, (size_t)NACC * NZ * 8);
    memset(turb, 0, (size_t)19 * NZ * 8);
    memset(vort, 0, (size_t)3 * NZ * 8);
    memset(ke, 0, (size_t)30 * NZ * 8);
#define S(n) (a + (size_t)(n) * NZ)
    for (k = 0; k < NZ; k++) { /* :1003-1058 */
        const double num = k < zdf ? nf : ((k < zdf + zdc - 4 && k >= zdf) ? nc : nf);
        for (n = 0; n < NACC; n++) S(n)[k] /= num;
    }
    for (k = 1; k < NZ - 1; k++) { /* :1201-1237 */
        const double siu = UTIN[k] = sqrt(S(A_UU)[k]), siv = VTIN[k] = sqrt(S(A_VV)[k]);
        const double siw = WTIN[k] = sqrt(S(A_WW)[k]), sip = DTIN[k] = sqrt(S(A_DD)[k]);
        KE[k] += (S(A_UU)[k] + S(A_VV)[k] + S(A_WW)[k]);
        S(A_WX)[k] = sqrt(S(A_WX)[k]); S(A_WY)[k] = sqrt(S(A_WY)[k]); S(A_WZ)[k] = sqrt(S(A_WZ)[k]);
        S(A_SKU)[k] /= (siu * siu * siu); S(A_SKV)[k] /= (siv * siv * siv); S(A_SKW)[k] /= (siw * siw * siw);
        S(A_SKP)[k] /= (sip * sip * sip);
        S(A_SKUW)[k] /= (siu * siu * siu * siw * siw * siw);
        S(A_SKVW)[k] /= (siv * siv * siv * siw * siw * siw);
        S(A_SKUV)[k] /= (siu * siu * siu * siv * siv * siv);
        S(A_KUU)[k] /= (siu * siu * siu * siu); S(A_KUV)[k] /= (siv * siv * siv * siv);
        S(A_KUW)[k] /= (siw * siw * siw * siw); S(A_KUP)[k] /= (sip * sip * sip * sip);
        S(A_KUUW)[k] /= (siu * siu * siu * siu * siw * siw * siw * siw);
        S(A_KUUV)[k] /= (siu * siu * siu * siu * siv * siv * siv * siv);
        S(A_KUVW)[k] /= (siv * siv * siv * siv * siw * siw * siw * siw);
        S(A_EPUU)[k] *= nu; S(A_EPVV)[k] *= nu; S(A_EPWW)[k] *= nu; S(A_EPUW)[k] *= (2. * nu);
        Epii[k] = (S(A_EPUU)[k]) + (S(A_EPVV)[k]) + (S(A_EPWW)[k]);
    }
    /* production and transport (:1238-1547); D1F, D1B, D2C, NUl, NUu are the tool's
       DDZ1F, DDZ1B, DDZ2C, DDZ2CNUl, DDZ2CNUu */
#define D1F(s, k) ((S(s)[(k) + 1]) - (S(s)[(k)]))
#define D1B(s, k) ((S(s)[(k)]) - (S(s)[(k) - 1]))
#define D2C(s, k) (((S(s)[(k) + 1]) - (S(s)[(k) - 1])) * 0.5)
#define NUL(s, k) ((1. / (1. + GR)) * ((1. / GR) * ((S(s)[(k) + 1]) - (S(s)[(k)])) + GR * ((S(s)[(k)]) - (S(s)[(k) - 1]))))
#define NUU(s, k) ((1. / (1. + GR)) * ((1. / GR) * ((S(s)[(k)]) - (S(s)[(k) - 1])) + GR * ((S(s)[(k) + 1]) - (S(s)[(k)]))))
#define TERMS(OP, m, m2, cs, half)                                                                                   \
    do {                                                                                                             \
        Tuur[q] = (half) * OP(A_UUW, q); Tvvr[q] = (half) * OP(A_VVW, q); Twwr[q] = (half) * OP(A_WWW, q);            \
        Tuwr[q] = (m) * OP(A_UWW, q); Tiipai[q] = (m) * OP(A_PW, q); Twwpai[q] = (m2) * OP(A_PW, q);                  \
        Tuwpai[q] = (m) * OP(A_PU, q);                                                                               \
        Tuus[q] = (cs) * OP(A_USUW, q); Tvvs[q] = (cs) * OP(A_VSVW, q); Twws[q] = (cs) * OP(A_WSWW, q);              \
        Tuws[q] = (cs) * OP(A_WSUW, q) + (cs) * OP(A_WSWW, q);                                                       \
    } while (0)
    {
        const int lo = 0, mid = zdf - 2, up = zdf + zdc - 4;
        for (k = 1; k < zdf; k++) {
            const int q = lo + k;
            Pii[q] = -(S(A_UW)[q]) * (dUdz[q]);
            Puu[q] = 2. * (Pii[q]); Pvv[q] = 0.; Pww[q] = 0.;
            Puw[q] = -(S(A_WW)[q]) * (dUdz[q]);
            Tuupai[q] = 0.; Tvvpai[q] = 0.;
            if (k == 1) TERMS(D1F, -1., -2., 2. * nu, -1.);
            else if (k == zdf - 1) TERMS(NUL, -1., -2., 2. * nu, -1.);
            else if (k == 2 || k == zdf - 2) TERMS(D2C, -1., -2., 2. * nu, -1.);
            else TERMS(D2C, -1., -2., 2. * nu, -0.5);
            Tiis[q] = 0.5 * (Tuus[q] + Tvvs[q] + Twws[q]);
            Tiir[q] = 0.5 * (Tuur[q] + Tvvr[q] + Twwr[q]);
        }
        for (k = 2; k < zdc - 2; k++) {
            const int q = mid + k;
            Pii[q] = -(S(A_UW)[q]) * (dUdz[q]);
            Puu[q] = 2. * (Pii[q]); Pvv[q] = 0.; Pww[q] = 0.;
            Puw[q] = -(S(A_WW)[q]) * (dUdz[q]);
            Tuupai[q] = 0.; Tvvpai[q] = 0.;
            if (k == 2 || k == 3 || k == zdc - 4 || k == zdc - 3) TERMS(D2C, -coef, -2. * coef, 2. * nu * coef, -coef);
            else TERMS(D2C, -coef, -2. * coef, coef * 2. * nu, -coef * 0.5);
            Tiis[q] = 0.5 * (Tuus[q] + Tvvs[q] + Twws[q]);
            Tiir[q] = 0.5 * (Tuur[q] + Tvvr[q] + Twwr[q]);
        }
        for (k = 0; k < zdf - 1; k++) {
            const int q = up + k;
            Pii[q] = -(S(A_UW)[q]) * (dUdz[q]);
            Puu[q] = 2. * (Pii[q]); Pvv[q] = 0.; Pww[q] = 0.;
            Puw[q] = -(S(A_WW)[q]) * (dUdz[q]);
            Tuupai[q] = 0.; Tvvpai[q] = 0.;
            if (k == 0) TERMS(NUU, -1., -2., 2. * nu, -1.);
            else if (k == zdf - 2) TERMS(D1B, -1., -2., 2. * nu, -1.);
            else if (k == 2 || k == zdf - 3) TERMS(D2C, -1., -2., 2. * nu, -1.);
            else TERMS(D2C, -1., -2., 2. * nu, -0.5);
            Tiis[q] = 0.5 * (Tuus[q] + Tvvs[q] + Twws[q]);
            Tiir[q] = 0.5 * (Tuur[q] + Tvvr[q] + Twwr[q]);
        }
    }
    /* the files' profiles (:1553-1636) */
    memcpy(turb, Um, (size_t)3 * NZ * 8);          /* UAVG VAVG WAVG */
    memcpy(turb + 6 * NZ, S(A_UW), NZ * 8);
    memcpy(turb + 7 * NZ, Um + 3 * NZ, NZ * 8);    /* DAVG */
    memcpy(turb + 9 * NZ, S(A_SKU), NZ * 8); memcpy(turb + 10 * NZ, S(A_SKV), NZ * 8);
    memcpy(turb + 11 * NZ, S(A_SKW), NZ * 8); memcpy(turb + 12 * NZ, S(A_SKUW), NZ * 8);
    memcpy(turb + 13 * NZ, S(A_SKP), NZ * 8);
    memcpy(turb + 14 * NZ, S(A_KUU), NZ * 8); memcpy(turb + 15 * NZ, S(A_KUV), NZ * 8);
    memcpy(turb + 16 * NZ, S(A_KUW), NZ * 8); memcpy(turb + 17 * NZ, S(A_KUUW), NZ * 8);
    memcpy(turb + 18 * NZ, S(A_KUP), NZ * 8);
    memcpy(vort, S(A_WX), NZ * 8); memcpy(vort + NZ, S(A_WY), NZ * 8); memcpy(vort + 2 * NZ, S(A_WZ), NZ * 8);
    memcpy(ke + 15 * NZ, S(A_PAIUU), NZ * 8); memcpy(ke + 16 * NZ, S(A_PAIVV), NZ * 8);
    memcpy(ke + 17 * NZ, S(A_PAIWW), NZ * 8); memcpy(ke + 18 * NZ, S(A_PAIUW), NZ * 8);
    memcpy(ke + 25 * NZ, S(A_EPUU), NZ * 8); memcpy(ke + 26 * NZ, S(A_EPVV), NZ * 8);
    memcpy(ke + 27 * NZ, S(A_EPWW), NZ * 8); memcpy(ke + 28 * NZ, S(A_EPUW), NZ * 8);
#undef S
    free(a);
}
