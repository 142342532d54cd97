// budget.hip -- per-snapshot turbulence statistics and kinetic-energy budget on the
// device (dnsg_budget2d).
//
// The reference computes them offline, one snapshot at a time, with its TKE-budget
// tool (TKE-budget/computations/budget2D.c:13-1736, reading.c, uv_fluc.c) from the
// vel-* and den-* files.  Here the same pass runs on the current state:
//   1. k_bud_fill    velocity and density of every node of the three grids in the
//                    tool's array layout, [X+4][Y][Z] with two periodic ghost planes
//                    on each side in x (reading.c:33-57), formed as vel_print does;
//   2. k_bud_layer   per-layer sums of u, v, w, rho/3 (budget2D.c:177-244), reduced
//                    in two stages in a fixed order;
//   3. k_bud_sub     the layer means subtracted in place, ghost planes included
//                    (:265-309), so every read below sees what the tool's arrays hold,
//                    also where a one-sided z stencil reaches into the next row;
//   4. k_bud_nodes   the 44 per-node products and derivative terms of :449-987,
//                    summed per layer in two stages;
//   5. the O(NZ) tail on the host: dU/dz, normalisation, production, transport,
//      dissipation (:311-444, :1003-1547).
// Each node's terms are formed exactly as the tool forms them; the sums over a
// layer are taken in another order, so the profiles agree with the tool's to
// rounding (1e-12 relative), not bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {

namespace {

// the per-node accumulators (budget2D.c:449-617), in this order in every table
enum {
    A_UU, A_UV, A_UW, A_VV, A_VW, A_WW, A_SKU, A_SKV, A_SKW, A_SKP, A_SKUW, A_SKUV, A_SKVW,
    A_KUU, A_KUV, A_KUW, A_KUP, A_KUUW, A_KUVW, A_KUUV, A_DD, A_UUW, A_VVW, A_WWW, A_UWW,
    A_PU, A_PV, A_PW, A_WX, A_WY, A_WZ, A_PAIUU, A_PAIVV, A_PAIWW, A_PAIUW,
    A_USUW, A_VSVW, A_WSWW, A_WSUW, A_USWW, A_EPUU, A_EPVV, A_EPWW, A_EPUW, NACC
};
constexpr int BT = 256;  // threads per workgroup

// one grid of the snapshot in the tool's layout
struct BudGrid {
    double *vel, *den;  // [X+4][Y][Z][3], [X+4][Y][Z]
    int X, Y, Z;
    int k0, k1;         // layers taking part in the statistics
    int base;           // profile index of layer 0
};

struct BudArgs {
    BudGrid g[3];       // coarse, lower band, upper band
    int GR;
    double coef;
    const double *mean; // [4][NZ]: U, V, W, rho/3
    int NZ;
};

__device__ __forceinline__ long bidx(const BudGrid &G, int i, int j, int k) {
    return ((long)i + 2) * G.Y * G.Z + (long)j * G.Z + k;  // the tool's Fb with its +2 plane offset
}
__device__ __forceinline__ double vat(const BudGrid &G, int i, int j, int k, int c) {
    return G.vel[bidx(G, i, j, k) * 3 + c];
}

// 1. velocity and density as vel_print.c:86-117 writes them, [X+2 pad][Y][Z] with
//    pad periodic ghost planes on each side in x (pad 0: the vel-* / den-* file
//    layout); the band wall layers come from the wall-layer copies [X][Y][19]
template <class S>
__global__ __launch_bounds__(BT) void k_bud_fill(const S *__restrict__ lat, const double *__restrict__ wall, int X,
                                                 int Y, int Z, int zlo, int kwall, int pad, double *__restrict__ vel,
                                                 double *__restrict__ den) {
    const long n = (long)(X + 2 * pad) * Y * Z;
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= n) return;
    const int k = (int)(id % Z), j = (int)((id / Z) % Y), p = (int)(id / ((long)Y * Z));
    const int x = (p - pad + X) % X;
    double f[Q];
    if (k == kwall) {
        const double *w = wall + ((size_t)x * Y + j) * Q;
        unroll<0, Q>([&](auto qq) { f[decltype(qq)::value] = w[decltype(qq)::value]; });
    } else {
        const size_t row = (size_t)(k - zlo) * Y + j;
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            f[q] = (double)load_s(lat + lat_at(q, row, x, X));
            if constexpr (sizeof(S) != 8) f[q] += WQ[q];
        });
    }
    double ux = 0., uy = 0., uz = 0., dens = 0.;
    unroll<0, Q>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        ux += (double)EX[q] * f[q];
        uy += (double)EY[q] * f[q];
        uz += (double)EZ[q] * f[q];
        dens += f[q];
    });
    if (dens > 0.) { ux /= dens; uy /= dens; uz /= dens; }
    else { ux = 0.; uy = 0.; uz = 0.; dens = 0.; }
    vel[id * 3] = ux;
    vel[id * 3 + 1] = uy;
    vel[id * 3 + 2] = uz;
    den[id] = dens;
}

// sum of W values over a workgroup into out[0..W) (fixed order: lanes, then waves)
template <int W>
__device__ __forceinline__ void wg_sum(double *v, double *out) {
    __shared__ double l[BT / 64][W];
#pragma unroll
    for (int c = 0; c < W; c++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_down(v[c], o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < W; c++) l[wv][c] = v[c];
    __syncthreads();
    for (int c = threadIdx.x; c < W; c += BT) {
        double s = l[0][c];
#pragma unroll
        for (int w = 1; w < BT / 64; w++) s += l[w][c];
        out[c] = s;
    }
}

// chunks of one layer: a workgroup takes BT consecutive nodes n = i*Y + j
__host__ __device__ inline int bud_chunks(const BudGrid &G) { return (G.X * G.Y + BT - 1) / BT; }

// 2. per-(layer, chunk) sums of u, v, w, rho/3 -> part[(k - k0) * chunks + chunk][4]
__global__ __launch_bounds__(BT) void k_bud_layer(BudGrid G, double *__restrict__ part) {
    const int nch = bud_chunks(G);
    const int k = G.k0 + (int)blockIdx.x / nch, ch = (int)blockIdx.x % nch;
    const int n = ch * BT + threadIdx.x;
    double v[4] = {0., 0., 0., 0.};
    if (n < G.X * G.Y) {
        const int i = n / G.Y, j = n % G.Y;
        const long e = bidx(G, i, j, k);
        v[0] = G.vel[e * 3];
        v[1] = G.vel[e * 3 + 1];
        v[2] = G.vel[e * 3 + 2];
        v[3] = (1. / 3.) * G.den[e];
    }
    wg_sum<4>(v, part + (size_t)blockIdx.x * 4);
}

// sums of nch consecutive partial records of w values, one thread per (segment, value)
__global__ __launch_bounds__(BT) void k_bud_segsum(const double *__restrict__ part, int nseg, int nch, int w,
                                                   double *__restrict__ out) {
    const int t = blockIdx.x * BT + threadIdx.x;
    if (t >= nseg * w) return;
    const int s = t / w, c = t % w;
    double a = 0.;
    for (int i = 0; i < nch; i++) a += part[((size_t)s * nch + i) * w + c];
    out[t] = a;
}

// 3. fluctuations in place, ghost planes included
__global__ __launch_bounds__(BT) void k_bud_sub(BudGrid G, const double *__restrict__ mean, int NZ) {
    const long n = (long)(G.X + 4) * G.Y * (G.k1 - G.k0);
    const long id = (long)blockIdx.x * BT + threadIdx.x;
    if (id >= n) return;
    const int nk = G.k1 - G.k0;
    const int k = G.k0 + (int)(id % nk);
    const long pj = id / nk;  // (p, j)
    double *v = G.vel + (pj * G.Z + k) * 3;
    v[0] -= mean[G.base + k];
    v[1] -= mean[NZ + G.base + k];
    v[2] -= mean[2 * NZ + G.base + k];
}

// z derivative of component c at (i, j, k) of grid g (budget2D.c:539-593 lower band,
// :720-748 coarse, :904-940 upper band)
__device__ __forceinline__ double bud_ddz(const BudArgs &A, int g, int i, int j, int k, int c) {
    const BudGrid &G = A.g[g];
    const int GR = A.GR;
    const double coef = A.coef;
    auto V = [&](int kk) { return vat(G, i, j, kk, c); };
    if (g == 1) {
        if (k == 1) {
            const double tmp = (2. / 3.) * (V(k + 3) - V(k + 1)) - (1. / 12.) * (V(k + 4) - V(k));
            return -(7. / 3.) * V(k) + 6. * V(k + 1) - 3. * V(k + 2) - (2. / 3.) * V(k + 3) + 3. * tmp;
        }
        if (k == G.Z - 1) {
            const double cc = vat(A.g[0], i / GR, j / GR, 2, c);
            return (coef * (cc - V(k)) + GR * (V(k) - V(k - 1))) * 1. / (1. + GR);
        }
        return 0.5 * (V(k + 1) - V(k - 1));
    }
    if (g == 0) {
        if (k == 2) return 0.5 * coef * (V(k + 1) - vat(A.g[1], i * GR, j * GR, A.g[1].Z - 1, c));
        if (k == 3 || k == G.Z - 4) return 0.5 * coef * (V(k + 1) - V(k - 1));
        if (k == G.Z - 3) return 0.5 * coef * (vat(A.g[2], i * GR, j * GR, 0, c) - V(k - 1));
        return 0.5 * coef * (V(k + 1) - V(k - 1));
    }
    if (k == 0) {
        const double cc = vat(A.g[0], i / GR, j / GR, A.g[0].Z - 3, c);
        return (GR * (V(k + 1) - V(k)) + coef * (V(k) - cc)) * 1. / (1. + GR);
    }
    if (k == G.Z - 2) {
        const double tmp = (2. / 3.) * (V(k - 1) - V(k - 3)) - (1. / 12.) * (V(k) - V(k - 4));
        return (7. / 3.) * V(k) - 6. * V(k - 1) + 3. * V(k - 2) + (2. / 3.) * V(k - 3) + 3. * tmp;
    }
    return 0.5 * (V(k + 1) - V(k - 1));
}

// 4. the per-node terms of grid g (budget2D.c:449-617, :636-797, :800-987), summed
//    per (layer, chunk) -> part[(k - k0) * chunks + chunk][NACC]
__global__ __launch_bounds__(BT) void k_bud_nodes(BudArgs A, int g, double *__restrict__ part) {
    const BudGrid &G = A.g[g];
    const int nch = bud_chunks(G);
    const int k = G.k0 + (int)blockIdx.x / nch, ch = (int)blockIdx.x % nch;
    const int n = ch * BT + threadIdx.x;
    double a[NACC];
#pragma unroll
    for (int c = 0; c < NACC; c++) a[c] = 0.;
    if (n < G.X * G.Y) {
        const int Y = G.Y, i = n / Y, j = n % Y;
        const long e = bidx(G, i, j, k);
        const double u = G.vel[e * 3], vv = G.vel[e * 3 + 1], w = G.vel[e * 3 + 2];
        const double dfl = (1. / 3.) * G.den[e] - A.mean[3 * A.NZ + G.base + k];
        const int jp = j == 0 ? Y - 1 : j - 1, jpp = j == 0 ? Y - 2 : (j == 1 ? Y - 1 : j - 2);
        const int jn = j == Y - 1 ? 0 : j + 1, jnn = j == Y - 1 ? 1 : (j == Y - 2 ? 0 : j + 2);
        double d[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double gx = (2. / 3.) * (vat(G, i + 1, j, k, c) - vat(G, i - 1, j, k, c)) -
                              (1. / 12.) * (vat(G, i + 2, j, k, c) - vat(G, i - 2, j, k, c));
            const double gy = (2. / 3.) * (vat(G, i, jn, k, c) - vat(G, i, jp, k, c)) -
                              (1. / 12.) * (vat(G, i, jnn, k, c) - vat(G, i, jpp, k, c));
            d[c][0] = g == 0 ? A.coef * gx : gx;
            d[c][1] = g == 0 ? A.coef * gy : gy;
            d[c][2] = bud_ddz(A, g, i, j, k, c);
        }
        const double dudx = d[0][0], dudy = d[0][1], dudz = d[0][2];
        const double dvdx = d[1][0], dvdy = d[1][1], dvdz = d[1][2];
        const double dwdx = d[2][0], dwdy = d[2][1], dwdz = d[2][2];
        const double ox = (dwdy - dvdz), oy = (dudz - dwdx), oz = (dvdx - dudy);
        a[A_UU] = u * u; a[A_UV] = u * vv; a[A_UW] = u * w; a[A_VV] = vv * vv; a[A_VW] = vv * w; a[A_WW] = w * w;
        a[A_SKU] = u * u * u; a[A_SKV] = vv * vv * vv; a[A_SKW] = w * w * w; a[A_SKP] = dfl * dfl * dfl;
        a[A_SKUW] = u * u * u * w * w * w; a[A_SKUV] = u * u * u * vv * vv * vv; a[A_SKVW] = vv * vv * vv * w * w * w;
        a[A_KUU] = u * u * u * u; a[A_KUV] = vv * vv * vv * vv; a[A_KUW] = w * w * w * w;
        a[A_KUP] = dfl * dfl * dfl * dfl;
        a[A_KUUW] = u * u * u * u * w * w * w * w; a[A_KUVW] = vv * vv * vv * vv * w * w * w * w;
        a[A_KUUV] = u * u * u * u * vv * vv * vv * vv;
        a[A_DD] = dfl * dfl;
        a[A_UUW] = u * u * w; a[A_VVW] = vv * vv * w; a[A_WWW] = w * w * w; a[A_UWW] = u * w * w;
        a[A_PU] = u * dfl; a[A_PV] = vv * dfl; a[A_PW] = w * dfl;
        a[A_WX] = ox * ox; a[A_WY] = oy * oy; a[A_WZ] = oz * oz;
        a[A_PAIUU] = dfl * dudx; a[A_PAIVV] = dfl * dvdy; a[A_PAIWW] = dfl * dwdz; a[A_PAIUW] = dfl * (dudz + dwdx);
        a[A_USUW] = 0.5 * u * (dudz + dwdx); a[A_VSVW] = 0.5 * vv * (dvdz + dwdy);
        a[A_WSWW] = 0.5 * w * (dwdz + dwdz); a[A_WSUW] = 0.5 * w * (dudz + dwdx); a[A_USWW] = 0.5 * u * (dwdz + dwdz);
        a[A_EPUU] = dudx * (dudx + dudx) + dudy * (dudy + dvdx) + dudz * (dudz + dwdx);
        a[A_EPVV] = dvdx * (dvdx + dudy) + dvdy * (dvdy + dvdy) + dvdz * (dvdz + dwdy);
        a[A_EPWW] = dwdx * (dwdx + dudz) + dwdy * (dwdy + dvdz) + dwdz * (dwdz + dwdz);
        a[A_EPUW] = dudx * (dwdx + dudz) + dudy * (dwdy + dvdz) + dudz * (dwdz + dwdz) + dwdx * (dudx + dudx) +
                    dwdy * (dudy + dvdx) + dwdz * (dudz + dwdx);
    }
    wg_sum<NACC>(a, part + (size_t)blockIdx.x * NACC);
}

// the lower band's u', v' at k = 4 (uv_fluc.c:25-33), [Xf][Yf]
__global__ __launch_bounds__(BT) void k_bud_plane(BudGrid G, int k, double *__restrict__ uo, double *__restrict__ vo) {
    const int n = blockIdx.x * BT + threadIdx.x;
    if (n >= G.X * G.Y) return;
    const long e = bidx(G, n / G.Y, n % G.Y, k);
    uo[n] = G.vel[e * 3];
    vo[n] = G.vel[e * 3 + 1];
}

// 5. dU/dz of the mean profile (budget2D.c:311-444)
void bud_dudz(const double *U, int zdf, int zdc, int GR, double *dUdz) {
    const double coef = (1. / ((double)GR));
    const int zd = zdf;
    for (int k = 1; k < zd; k++) {
        const double *a = U;
        if (k == 1) {
            const double tmp = (2. / 3.) * (a[k + 3] - a[k + 1]) - (1. / 12.) * (a[k + 4] - a[k]);
            dUdz[k] = -(7. / 3.) * a[k] + 6. * a[k + 1] - 3. * a[k + 2] - (2. / 3.) * a[k + 3] + 3. * tmp;
        } else if (k == zd - 1) {
            dUdz[k] = (coef * (a[k + 1] - a[k]) + GR * (a[k] - a[k - 1])) * 1. / (1. + GR);
        } else if (k == 2 || k == zd - 2) {
            dUdz[k] = 0.5 * (a[k + 1] - a[k - 1]);
        } else {
            dUdz[k] = (2. / 3.) * (a[k + 1] - a[k - 1]) - (1. / 12.) * (a[k + 2] - a[k - 2]);
        }
    }
    for (int k = 2; k < zdc - 2; k++) {
        const double *a = U + zdf - 2;
        if (k == 2 || k == 3 || k == zdc - 4 || k == zdc - 3) dUdz[zdf - 2 + k] = 0.5 * coef * (a[k + 1] - a[k - 1]);
        else dUdz[zdf - 2 + k] = coef * ((2. / 3.) * (a[k + 1] - a[k - 1]) - (1. / 12.) * (a[k + 2] - a[k - 2]));
    }
    double *o = dUdz + zdf + zdc - 4;
    const double *a = U + zdf + zdc - 4;
    for (int k = 0; k < zd - 1; k++) {
        if (k == 0) {
            o[k] = (GR * (a[k + 1] - a[k]) + coef * (a[k] - a[k - 1])) * 1. / (1. + GR);
        } else if (k == zd - 2) {
            const double tmp = (2. / 3.) * (a[k - 1] - a[k - 3]) - (1. / 12.) * (a[k] - a[k - 4]);
            o[k] = (7. / 3.) * a[k] - 6. * a[k - 1] + 3. * a[k - 2] + (2. / 3.) * a[k - 3] + 3. * tmp;
        } else if (k == 1 || k == zd - 3) {
            o[k] = 0.5 * (a[k + 1] - a[k - 1]);
        } else {
            o[k] = (2. / 3.) * (a[k + 1] - a[k - 1]) - (1. / 12.) * (a[k + 2] - a[k - 2]);
        }
    }
}

// 5. normalisation and the derived profiles (budget2D.c:1003-1547)
void bud_finish(const dnsg_params &p, const double *Um, const double *dUdz, std::vector<double> &a, double *turb,
                double *vort, double *ke) {
    const int GR = p.gratio, zdf = p.zds + 1, zdc = p.zdv;
    const int NZ = 2 * zdf + (zdc - 4), NZs = 2 * zdf + GR * (zdc - 4) + GR - 1;
    const double tauf = (3. / p.Rebs) * p.ub * (NZs - 2.) + 0.5;  // :55-58
    const double nu = (1. / 6.) * (2. * tauf - 1.), coef = (1. / ((double)GR));
    const double nf = (double)(GR * p.xdv) * (GR * p.ydv), nc = (double)p.xdv * p.ydv;
    auto S = [&](int n) { return a.data() + (size_t)n * NZ; };
    std::memset(turb, 0, (size_t)19 * NZ * 8);
    std::memset(vort, 0, (size_t)3 * NZ * 8);
    std::memset(ke, 0, (size_t)30 * NZ * 8);
    double *UTIN = turb + 3 * NZ, *VTIN = turb + 4 * NZ, *WTIN = turb + 5 * NZ, *DTIN = turb + 8 * NZ;
    double *Pii = ke, *Puu = ke + NZ, *Pvv = ke + 2 * NZ, *Pww = ke + 3 * NZ, *Puw = ke + 4 * NZ;
    double *Tiir = ke + 5 * NZ, *Tuur = ke + 6 * NZ, *Tvvr = ke + 7 * NZ, *Twwr = ke + 8 * NZ, *Tuwr = ke + 9 * NZ;
    double *Tiipai = ke + 10 * NZ, *Tuupai = ke + 11 * NZ, *Tvvpai = ke + 12 * NZ, *Twwpai = ke + 13 * NZ;
    double *Tuwpai = ke + 14 * NZ, *Tiis = ke + 19 * NZ, *Tuus = ke + 20 * NZ, *Tvvs = ke + 21 * NZ;
    double *Twws = ke + 22 * NZ, *Tuws = ke + 23 * NZ, *Epii = ke + 24 * NZ, *KE = ke + 29 * NZ;
    for (int k = 0; k < NZ; k++) {  // :1003-1058
        const double num = k < zdf ? nf : ((k < zdf + zdc - 4 && k >= zdf) ? nc : nf);
        for (int n = 0; n < NACC; n++) S(n)[k] /= num;
    }
    for (int k = 1; k < NZ - 1; k++) {  // :1201-1237
        const double siu = UTIN[k] = std::sqrt(S(A_UU)[k]), siv = VTIN[k] = std::sqrt(S(A_VV)[k]);
        const double siw = WTIN[k] = std::sqrt(S(A_WW)[k]), sip = DTIN[k] = std::sqrt(S(A_DD)[k]);
        KE[k] += (S(A_UU)[k] + S(A_VV)[k] + S(A_WW)[k]);
        S(A_WX)[k] = std::sqrt(S(A_WX)[k]);
        S(A_WY)[k] = std::sqrt(S(A_WY)[k]);
        S(A_WZ)[k] = std::sqrt(S(A_WZ)[k]);
        S(A_SKU)[k] /= (siu * siu * siu);
        S(A_SKV)[k] /= (siv * siv * siv);
        S(A_SKW)[k] /= (siw * siw * siw);
        S(A_SKP)[k] /= (sip * sip * sip);
        S(A_SKUW)[k] /= (siu * siu * siu * siw * siw * siw);
        S(A_SKVW)[k] /= (siv * siv * siv * siw * siw * siw);
        S(A_SKUV)[k] /= (siu * siu * siu * siv * siv * siv);
        S(A_KUU)[k] /= (siu * siu * siu * siu);
        S(A_KUV)[k] /= (siv * siv * siv * siv);
        S(A_KUW)[k] /= (siw * siw * siw * siw);
        S(A_KUP)[k] /= (sip * sip * sip * sip);
        S(A_KUUW)[k] /= (siu * siu * siu * siu * siw * siw * siw * siw);
        S(A_KUUV)[k] /= (siu * siu * siu * siu * siv * siv * siv * siv);
        S(A_KUVW)[k] /= (siv * siv * siv * siv * siw * siw * siw * siw);
        S(A_EPUU)[k] *= nu;
        S(A_EPVV)[k] *= nu;
        S(A_EPWW)[k] *= nu;
        S(A_EPUW)[k] *= (2. * nu);
        Epii[k] = (S(A_EPUU)[k]) + (S(A_EPVV)[k]) + (S(A_EPWW)[k]);
    }
    // production and transport (:1238-1547); the z differences of the tool's
    // DDZ1F, DDZ1B, DDZ2C, DDZ2CNUl and DDZ2CNUu
    enum { D1F, D1B, D2C, NUL, NUU };
    auto D = [&](int op, int s, int k) {
        const double *x = S(s);
        switch (op) {
        case D1F: return (x[k + 1]) - (x[k]);
        case D1B: return (x[k]) - (x[k - 1]);
        case D2C: return ((x[k + 1]) - (x[k - 1])) * 0.5;
        case NUL: return (1. / (1. + GR)) * ((1. / GR) * ((x[k + 1]) - (x[k])) + GR * ((x[k]) - (x[k - 1])));
        default: return (1. / (1. + GR)) * ((1. / GR) * ((x[k]) - (x[k - 1])) + GR * ((x[k + 1]) - (x[k])));
        }
    };
    auto terms = [&](int q, int op, double m, double m2, double cs, double half) {
        Tuur[q] = half * D(op, A_UUW, q);
        Tvvr[q] = half * D(op, A_VVW, q);
        Twwr[q] = half * D(op, A_WWW, q);
        Tuwr[q] = m * D(op, A_UWW, q);
        Tiipai[q] = m * D(op, A_PW, q);
        Twwpai[q] = m2 * D(op, A_PW, q);
        Tuwpai[q] = m * D(op, A_PU, q);
        Tuus[q] = cs * D(op, A_USUW, q);
        Tvvs[q] = cs * D(op, A_VSVW, q);
        Twws[q] = cs * D(op, A_WSWW, q);
        Tuws[q] = cs * D(op, A_WSUW, q) + cs * D(op, A_WSWW, q);
    };
    auto common = [&](int q) {
        Pii[q] = -(S(A_UW)[q]) * (dUdz[q]);
        Puu[q] = 2. * (Pii[q]);
        Pvv[q] = 0.;
        Pww[q] = 0.;
        Puw[q] = -(S(A_WW)[q]) * (dUdz[q]);
        Tuupai[q] = 0.;
        Tvvpai[q] = 0.;
    };
    auto sums = [&](int q) {
        Tiis[q] = 0.5 * (Tuus[q] + Tvvs[q] + Twws[q]);
        Tiir[q] = 0.5 * (Tuur[q] + Tvvr[q] + Twwr[q]);
    };
    for (int k = 1; k < zdf; k++) {
        const int q = k;
        common(q);
        if (k == 1) terms(q, D1F, -1., -2., 2. * nu, -1.);
        else if (k == zdf - 1) terms(q, NUL, -1., -2., 2. * nu, -1.);
        else if (k == 2 || k == zdf - 2) terms(q, D2C, -1., -2., 2. * nu, -1.);
        else terms(q, D2C, -1., -2., 2. * nu, -0.5);
        sums(q);
    }
    for (int k = 2; k < zdc - 2; k++) {
        const int q = zdf - 2 + k;
        common(q);
        if (k == 2 || k == 3 || k == zdc - 4 || k == zdc - 3) terms(q, D2C, -coef, -2. * coef, 2. * nu * coef, -coef);
        else terms(q, D2C, -coef, -2. * coef, coef * 2. * nu, -coef * 0.5);
        sums(q);
    }
    for (int k = 0; k < zdf - 1; k++) {
        const int q = zdf + zdc - 4 + k;
        common(q);
        if (k == 0) terms(q, NUU, -1., -2., 2. * nu, -1.);
        else if (k == zdf - 2) terms(q, D1B, -1., -2., 2. * nu, -1.);
        else if (k == 2 || k == zdf - 3) terms(q, D2C, -1., -2., 2. * nu, -1.);
        else terms(q, D2C, -1., -2., 2. * nu, -0.5);
        sums(q);
    }
    // the files' profiles (:1553-1636)
    std::memcpy(turb, Um, (size_t)3 * NZ * 8);
    std::memcpy(turb + 6 * NZ, S(A_UW), NZ * 8);
    std::memcpy(turb + 7 * NZ, Um + 3 * NZ, NZ * 8);
    const int tsrc[10] = {A_SKU, A_SKV, A_SKW, A_SKUW, A_SKP, A_KUU, A_KUV, A_KUW, A_KUUW, A_KUP};
    for (int t = 0; t < 10; t++) std::memcpy(turb + (9 + t) * NZ, S(tsrc[t]), NZ * 8);
    for (int t = 0; t < 3; t++) std::memcpy(vort + t * NZ, S(A_WX + t), NZ * 8);
    for (int t = 0; t < 4; t++) std::memcpy(ke + (15 + t) * NZ, S(A_PAIUU + t), NZ * 8);
    for (int t = 0; t < 4; t++) std::memcpy(ke + (25 + t) * NZ, S(A_EPUU + t), NZ * 8);
}

}  // namespace

// velocity and density of grid g (0 coarse, 1 lower, 2 upper) from the current
// state into [X + 2 pad][Y][Z][3] and [X + 2 pad][Y][Z]; needs fwall
hipError_t launch_snapshot(dnsg_ctx *h, int g, int pad, double *vel, double *den) {
    const dnsg_dims &d = h->d;
    const Grid *lat = g == 0 ? &h->coarse : (g == 1 ? &h->lower : &h->upper);
    const int X = g == 0 ? d.Xc : d.Xf, Y = g == 0 ? d.Yc : d.Yf, Z = g == 0 ? d.zdc : d.zdf;
    const int zlo = g == 1 ? 1 : 0, kwall = g == 0 ? -1 : (g == 1 ? 0 : d.zdf - 1);
    const double *wall = h->fwall + (g == 2 ? (size_t)d.Xf * d.Yf * 19 : 0);
    const long n = (long)(X + 2 * pad) * Y * Z;
    const dim3 grd((unsigned)((n + BT - 1) / BT)), blk(BT);
    if (h->p.precision == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_bud_fill<double>, grd, blk, 0, h->stream, (const double *)lat->c(), wall, X, Y, Z, zlo,
                           kwall, pad, vel, den);
    else if (h->p.precision == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_bud_fill<float>, grd, blk, 0, h->stream, (const float *)lat->c(), wall, X, Y, Z, zlo,
                           kwall, pad, vel, den);
    else
        hipLaunchKernelGGL(k_bud_fill<__half>, grd, blk, 0, h->stream, (const __half *)lat->c(), wall, X, Y, Z, zlo,
                           kwall, pad, vel, den);
    return hipGetLastError();
}

}  // namespace dnsg

using namespace dnsg;

#define BCHK(h, expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            rc = e_ == hipErrorOutOfMemory ? DNSG_ERR_OOM : DNSG_ERR_HIP;                   \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)

extern "C" int dnsg_budget2d(dnsg_t *h, double *turb, double *vort, double *ke, double *ufl, double *vfl) {
    if (!h || !h->stream || !turb || !vort || !ke || (!ufl) != (!vfl)) return DNSG_ERR_ARG;
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    const dnsg_dims &d = h->d;
    if (d.Yc < 4 || d.zdc < 6) {  // the tool's periodic y stencil and its coarse z branches
        h->err = "the budget needs ydv >= 4 and zdv >= 6";
        return DNSG_ERR_ARG;
    }
    int rc = ensure_fwall(h);
    if (rc) return rc;
    const int zdf = d.zdf, zdc = d.zdc, NZ = 2 * zdf + (zdc - 4), GR = d.GR;
    BudArgs A{};
    A.GR = GR;
    A.coef = (1. / ((double)GR));
    A.NZ = NZ;
    const int X[3] = {d.Xc, d.Xf, d.Xf}, Y[3] = {d.Yc, d.Yf, d.Yf}, Z[3] = {zdc, zdf, zdf};
    const int k0[3] = {2, 1, 0}, k1[3] = {zdc - 2, zdf, zdf - 1};    // node passes
    const int m0[3] = {2, 0, 0}, m1[3] = {zdc - 2, zdf, zdf};        // means and subtraction
    const int base[3] = {zdf - 2, 0, zdf + zdc - 4};
    size_t npart = 0;
    std::vector<double> mean(4 * (size_t)NZ, 0.), acc((size_t)NACC * NZ, 0.), dUdz(NZ, 0.), lay;
    double *dmean = nullptr, *dpart = nullptr, *dsum = nullptr, *dpl = nullptr;
    for (int g = 0; g < 3; g++) {
        A.g[g] = BudGrid{nullptr, nullptr, X[g], Y[g], Z[g], m0[g], m1[g], base[g]};
        npart = std::max(npart, (size_t)bud_chunks(A.g[g]) * Z[g] * NACC);
    }
    for (int g = 0; g < 3; g++) {
        const size_t nn = (size_t)(X[g] + 4) * Y[g] * Z[g];
        BCHK(h, hipMalloc(&A.g[g].vel, nn * 3 * 8));
        BCHK(h, hipMalloc(&A.g[g].den, nn * 8));
    }
    BCHK(h, hipMalloc(&dmean, 4 * (size_t)NZ * 8));
    BCHK(h, hipMalloc(&dpart, npart * 8));
    BCHK(h, hipMalloc(&dsum, (size_t)NACC * NZ * 8));
    // 1. snapshot arrays
    for (int g = 0; g < 3; g++) {
        BCHK(h, launch_snapshot(h, g, 2, A.g[g].vel, A.g[g].den));
    }
    // 2. layer means
    for (int g = 0; g < 3; g++) {
        const BudGrid &G = A.g[g];
        const int nch = bud_chunks(G), nk = G.k1 - G.k0;
        hipLaunchKernelGGL(k_bud_layer, dim3((unsigned)(nk * nch)), dim3(BT), 0, h->stream, G, dpart);
        hipLaunchKernelGGL(k_bud_segsum, dim3((unsigned)((nk * 4 + BT - 1) / BT)), dim3(BT), 0, h->stream,
                           (const double *)dpart, nk, nch, 4, dsum);
        BCHK(h, hipGetLastError());
        lay.resize((size_t)nk * 4);
        BCHK(h, hipMemcpyAsync(lay.data(), dsum, lay.size() * 8, hipMemcpyDeviceToHost, h->stream));
        BCHK(h, hipStreamSynchronize(h->stream));
        for (int k = G.k0; k < G.k1; k++)
            for (int c = 0; c < 4; c++)
                mean[(size_t)c * NZ + G.base + k] = lay[(size_t)(k - G.k0) * 4 + c] / (double)(G.X * G.Y);
    }
    BCHK(h, hipMemcpyAsync(dmean, mean.data(), mean.size() * 8, hipMemcpyHostToDevice, h->stream));
    // 3. fluctuations
    for (int g = 0; g < 3; g++) {
        const BudGrid &G = A.g[g];
        const long n = (long)(G.X + 4) * G.Y * (G.k1 - G.k0);
        hipLaunchKernelGGL(k_bud_sub, dim3((unsigned)((n + BT - 1) / BT)), dim3(BT), 0, h->stream, G,
                           (const double *)dmean, NZ);
        BCHK(h, hipGetLastError());
    }
    // 4. per-node terms over the layers of the node passes
    A.mean = dmean;
    for (int g = 0; g < 3; g++) {
        A.g[g].k0 = k0[g];
        A.g[g].k1 = k1[g];
    }
    for (int g = 0; g < 3; g++) {
        const BudGrid &G = A.g[g];
        const int nch = bud_chunks(G), nk = G.k1 - G.k0;
        hipLaunchKernelGGL(k_bud_nodes, dim3((unsigned)(nk * nch)), dim3(BT), 0, h->stream, A, g, dpart);
        hipLaunchKernelGGL(k_bud_segsum, dim3((unsigned)((nk * NACC + BT - 1) / BT)), dim3(BT), 0, h->stream,
                           (const double *)dpart, nk, nch, (int)NACC, dsum);
        BCHK(h, hipGetLastError());
        lay.resize((size_t)nk * NACC);
        BCHK(h, hipMemcpyAsync(lay.data(), dsum, lay.size() * 8, hipMemcpyDeviceToHost, h->stream));
        BCHK(h, hipStreamSynchronize(h->stream));
        for (int k = G.k0; k < G.k1; k++)
            for (int c = 0; c < NACC; c++) acc[(size_t)c * NZ + G.base + k] = lay[(size_t)(k - G.k0) * NACC + c];
    }
    if (ufl) {
        const BudGrid &G = A.g[1];
        const size_t np = (size_t)G.X * G.Y;
        BCHK(h, hipMalloc(&dpl, 2 * np * 8));
        hipLaunchKernelGGL(k_bud_plane, dim3((unsigned)((np + BT - 1) / BT)), dim3(BT), 0, h->stream, G, 4, dpl,
                           dpl + np);
        BCHK(h, hipGetLastError());
        BCHK(h, hipMemcpyAsync(ufl, dpl, np * 8, hipMemcpyDeviceToHost, h->stream));
        BCHK(h, hipMemcpyAsync(vfl, dpl + np, np * 8, hipMemcpyDeviceToHost, h->stream));
        BCHK(h, hipStreamSynchronize(h->stream));
    }
    // 5. the profile tail
    bud_dudz(mean.data(), zdf, zdc, GR, dUdz.data());
    bud_finish(h->p, mean.data(), dUdz.data(), acc, turb, vort, ke);
done:
    if (rc) (void)hipStreamSynchronize(h->stream);
    for (int g = 0; g < 3; g++) {
        if (A.g[g].vel) (void)hipFree(A.g[g].vel);
        if (A.g[g].den) (void)hipFree(A.g[g].den);
    }
    if (dmean) (void)hipFree(dmean);
    if (dpart) (void)hipFree(dpart);
    if (dsum) (void)hipFree(dsum);
    if (dpl) (void)hipFree(dpl);
    return rc;
}

// ---- time averages of the per-snapshot profiles ----------------------------------------
// The reference averages the per-snapshot files afterwards with its serial
// post-processor: time_stat.c:132-210 (turb-fld-2d: the mean of every profile except
// UTIN, VTIN, WTIN and DTIN, whose squares it averages), vorticity.c:73-98 (the
// squares of the three vorticity rms) and Ebudget.c:184-300 (the mean of every budget
// term).  Here the same sums are folded at each snapshot, in FP64, in the handle: the
// per-snapshot pass over the lattices is the device's work; the fold adds 52 profiles
// of NZ values per snapshot.
namespace {
constexpr int BUD_ROWS = 19 + 3 + 30;
// turb rows averaged as squares: UTIN VTIN WTIN (3..5) and DTIN (8); every vort row
inline bool bud_squared(int row) { return row == 3 || row == 4 || row == 5 || row == 8 || (row >= 19 && row < 22); }
}  // namespace

extern "C" int dnsg_budget2d_accumulate(dnsg_t *h, double *turb, double *vort, double *ke) {
    if (!h || !h->stream) return DNSG_ERR_ARG;
    const int NZ = h->d.NZ;
    std::vector<double> prof((size_t)BUD_ROWS * NZ);
    double *t = prof.data(), *v = t + 19 * (size_t)NZ, *k = v + 3 * (size_t)NZ;
    int rc = dnsg_budget2d(h, t, v, k, nullptr, nullptr);
    if (rc) return rc;
    if (h->bud_acc.size() != prof.size()) {
        h->bud_acc.assign(prof.size(), 0.);
        h->bud_count = 0;
    }
    for (int r = 0; r < BUD_ROWS; r++)
        for (int z = 0; z < NZ; z++) {
            const double x = prof[(size_t)r * NZ + z];
            h->bud_acc[(size_t)r * NZ + z] += bud_squared(r) ? x * x : x;
        }
    h->bud_count++;
    if (turb) memcpy(turb, t, 19 * (size_t)NZ * 8);
    if (vort) memcpy(vort, v, 3 * (size_t)NZ * 8);
    if (ke) memcpy(ke, k, 30 * (size_t)NZ * 8);
    return DNSG_OK;
}

extern "C" int dnsg_budget2d_average(dnsg_t *h, double *turb, double *vort, double *ke, long *count) {
    if (!h || !turb || !vort || !ke) return DNSG_ERR_ARG;
    const int NZ = h->d.NZ;
    if (count) *count = h->bud_count;
    if (h->bud_count == 0) {
        h->err = "no snapshot folded: call dnsg_budget2d_accumulate first";
        return DNSG_ERR_STATE;
    }
    double *dst[3] = {turb, vort, ke};
    const int rows[3] = {19, 3, 30};
    int r0 = 0;
    for (int g = 0; g < 3; g++) {
        for (size_t i = 0; i < (size_t)rows[g] * NZ; i++) dst[g][i] = h->bud_acc[(size_t)r0 * NZ + i] / h->bud_count;
        r0 += rows[g];
    }
    return DNSG_OK;
}

extern "C" int dnsg_budget2d_reset(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    h->bud_acc.clear();
    h->bud_count = 0;
    return DNSG_OK;
}
