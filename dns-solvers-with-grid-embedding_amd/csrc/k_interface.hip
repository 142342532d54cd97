// k_interface.hip -- the grid-embedding interface between the coarse grid and
// the two wall bands.
//
//   sample     coarse k = 1 / zdc-2 after streaming -> (rho, u, f - feq) planes
//              (xcomputations.c:558-625)
//   hermite    cubic Hermite in time of the three newest planes
//              (tinterpolate_hrmt.c:16-55)
//   deriv      4th-order central differences d/dx, d/dy, d2/dxdy
//              (derivative.c:17-89)
//   ctfd       bicubic patch per coarse cell (bcucof_opt.c:16-62), evaluated at
//              the GR x GR fine columns, f = feq + tau_f/(GR tau_c) f_neq on the
//              band interface layer (ctfdtransfer_opt.c:17-175)
//   ftcd       fine -> coarse injection at coincident nodes,
//              f = feq + GR tau_c/tau_f f_neq (ftcdtransfer.c:16-118)
// All planes are periodic in x and y, FP64, laid out [value][Yc][Xc].
// In the FP64 tier every expression keeps the reference's operation order, so
// the results are bitwise those of the reference.  In the shifted tiers the
// lattices hold g = f - w; planes stay FP64 and f_neq = g - (feq - w).
#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {

constexpr int NV = 23;  // rho, u_x, u_y, u_z, f_neq[19]

// ---- sampling --------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(256) void k_sample(const S *__restrict__ cs, int nx, int ny, int nz, int k_lo, int k_up,
                                                double *__restrict__ plo, double *__restrict__ pup) {
    const int n = nx * ny;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int k = blockIdx.y == 0 ? k_lo : k_up;
    double *pl = blockIdx.y == 0 ? plo : pup;
    const size_t qs = (size_t)nx * ny * nz, base = (size_t)k * n + id;
    double f[Q];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (sizeof(S) == 8) f[q] = cs[q * qs + base];
        else f[q] = (double)load_s(cs + q * qs + base);
    });
    if constexpr (sizeof(S) == 8) {
        const Macro64 m = macro64(f);
        pl[0 * n + id] = m.dens;
        pl[1 * n + id] = m.u0;
        pl[2 * n + id] = m.u1;
        pl[3 * n + id] = m.u2;
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            pl[(4 + q) * n + id] = f[q] - feq64<q>(m.dens, m.u0, m.u1, m.u2, m.v);
        });
    } else {
        const double drho = zeroth_moment(f), dens = 1. + drho;
        const double u0 = first_moment<0>(f) / dens, u1 = first_moment<1>(f) / dens, u2 = first_moment<2>(f) / dens;
        const double v = u0 * u0 + u1 * u1 + u2 * u2;
        pl[0 * n + id] = dens;
        pl[1 * n + id] = u0;
        pl[2 * n + id] = u1;
        pl[3 * n + id] = u2;
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const double eu = edot<q>(u0, u1, u2);
            const double geq = WQ[q] * (drho + dens * (3. * eu + 4.5 * eu * eu - 1.5 * v));
            pl[(4 + q) * n + id] = f[q] - geq;
        });
    }
}

hipError_t launch_sample(hipStream_t s, dnsg_ctx *h, const void *coarse_state, double *lo, double *up) {
    const Grid &g = h->coarse;
    const int n = g.nx * g.ny;
    dim3 grd((n + 255) / 256, 2), blk(256);
    const int prec = h->p.precision;
    if (prec == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_sample<double>, grd, blk, 0, s, (const double *)coarse_state, g.nx, g.ny, g.nz, 1,
                           g.nz - 2, lo, up);
    else if (prec == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_sample<float>, grd, blk, 0, s, (const float *)coarse_state, g.nx, g.ny, g.nz, 1,
                           g.nz - 2, lo, up);
    else
        hipLaunchKernelGGL(k_sample<__half>, grd, blk, 0, s, (const __half *)coarse_state, g.nx, g.ny, g.nz, 1,
                           g.nz - 2, lo, up);
    return hipGetLastError();
}

// ---- Hermite in time -----------------------------------------------------------------
struct HermiteW {
    double h00, h10, h01, h11;
};

struct Levels {
    const double *ln, *lp, *lpp;
    double *out;
};

__global__ __launch_bounds__(256) void k_hermite(Levels L0, Levels L1, int n, HermiteW H) {
    const Levels &L = blockIdx.y == 0 ? L0 : L1;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const double ln = L.ln[id], lp = L.lp[id], lpp = L.lpp[id];
    const double c1 = lp;
    const double c2 = (ln - lpp) * 0.5;
    const double c3 = ln;
    const double c4 = (3. * ln - 4. * lp + lpp) * 0.5;
    L.out[id] = H.h00 * c1 + H.h10 * c2 + H.h01 * c3 + H.h11 * c4;
}

// ---- 4th-order derivatives (periodic) -------------------------------------------------
__device__ __forceinline__ double d4(double fm2, double fm1, double fp1, double fp2) {
    return (2. / 3.) * (fp1 - fm1) - (1. / 12.) * (fp2 - fm2);
}

__global__ __launch_bounds__(256) void k_deriv12(double *s_lo, double *s_up, int nx, int ny) {
    double *sl = blockIdx.y == 0 ? s_lo : s_up;
    const int n = nx * ny;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * NV) return;
    const int v = id / n, r = id % n, x = r % nx, y = r / nx;
    const double *s0 = sl + (size_t)v * n;
    auto X = [&](int d) { int t = x + d; t += (t < 0) ? nx : 0; t -= (t >= nx) ? nx : 0; return t; };
    auto Y = [&](int d) { int t = y + d; t += (t < 0) ? ny : 0; t -= (t >= ny) ? ny : 0; return t; };
    sl[(size_t)(1 * NV + v) * n + r] = d4(s0[y * nx + X(-2)], s0[y * nx + X(-1)], s0[y * nx + X(1)], s0[y * nx + X(2)]);
    sl[(size_t)(2 * NV + v) * n + r] =
        d4(s0[Y(-2) * nx + x], s0[Y(-1) * nx + x], s0[Y(1) * nx + x], s0[Y(2) * nx + x]);
}

__global__ __launch_bounds__(256) void k_deriv3(double *s_lo, double *s_up, int nx, int ny) {
    double *sl = blockIdx.y == 0 ? s_lo : s_up;
    const int n = nx * ny;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * NV) return;
    const int v = id / n, r = id % n, x = r % nx, y = r / nx;
    const double *s2 = sl + (size_t)(2 * NV + v) * n;
    auto X = [&](int d) { int t = x + d; t += (t < 0) ? nx : 0; t -= (t >= nx) ? nx : 0; return t; };
    sl[(size_t)(3 * NV + v) * n + r] = d4(s2[y * nx + X(-2)], s2[y * nx + X(-1)], s2[y * nx + X(1)], s2[y * nx + X(2)]);
}

// ---- bicubic coefficients and the fine interface layer ------------------------------
// the bicubic interpolation matrix of bcucof_opt.c:18-35 (Press et al., bcucof)
constexpr int WT[16][16] = {
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

// c = WT x, summed in k order from 0 with the zero weights dropped
__device__ __forceinline__ void bicubic_coef(const double *xv, double *c) {
    unroll<0, 16>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        double xx = 0.0;
        unroll<0, 16>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (WT[i][k] != 0) xx = xx + (double)WT[i][k] * xv[k];
        });
        c[i] = xx;
    });
}

// Horner form of ctfdtransfer_opt.c:59-62
__device__ __forceinline__ double bicubic_eval(const double *c, double tt, double uu) {
    double r = 0.;
    for (int nn = 3; nn > -1; nn--) r = tt * r + ((c[nn * 4 + 3] * uu + c[nn * 4 + 2]) * uu + c[nn * 4 + 1]) * uu + c[nn * 4 + 0];
    return r;
}

__device__ __forceinline__ void gather16(const double *sl, int n, int v, int p00, int p10, int p11, int p01,
                                         double *xv) {
    const double *s0 = sl + (size_t)(0 * NV + v) * n, *s1 = sl + (size_t)(1 * NV + v) * n;
    const double *s2 = sl + (size_t)(2 * NV + v) * n, *s3 = sl + (size_t)(3 * NV + v) * n;
    xv[0] = s0[p00]; xv[4] = s1[p00]; xv[8] = s2[p00];  xv[12] = s3[p00];
    xv[1] = s0[p10]; xv[5] = s1[p10]; xv[9] = s2[p10];  xv[13] = s3[p10];
    xv[2] = s0[p11]; xv[6] = s1[p11]; xv[10] = s2[p11]; xv[14] = s3[p11];
    xv[3] = s0[p01]; xv[7] = s1[p01]; xv[11] = s2[p01]; xv[15] = s3[p01];
}

struct CtfdBand {
    const double *slots;
    void *band;   // band post-stream lattice
    int layer;    // stored interface layer
};

template <class S, int GR>
__global__ __launch_bounds__(128) void k_ctfd(CtfdBand B0, CtfdBand B1, int Xc, int Yc, int nzb, double rescale) {
    const CtfdBand &B = blockIdx.y == 0 ? B0 : B1;
    const int n = Xc * Yc;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ic = id % Xc, jc = id / Xc;
    const int in = (ic + 1 == Xc) ? 0 : ic + 1, jn = (jc + 1 == Yc) ? 0 : jc + 1;
    const int p00 = jc * Xc + ic, p10 = jc * Xc + in, p11 = jn * Xc + in, p01 = jn * Xc + ic;
    const int Xf = GR * Xc, Yf = GR * Yc;
    const size_t qs = (size_t)Xf * Yf * nzb;
    S *band = static_cast<S *>(B.band);
    const size_t lay0 = (size_t)B.layer * Yf * Xf;
    constexpr int NP = GR * GR;
    double dens[NP], u0[NP], u1[NP], u2[NP], vv[NP];
    double xv[16], c[16];
    // rho and u at the fine points
    for (int v = 0; v < 4; v++) {
        gather16(B.slots, n, v, p00, p10, p11, p01, xv);
        bicubic_coef(xv, c);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int ii = p / GR, jj = p % GR;
            const double tt = ((double)ii) / ((double)GR), uu = ((double)jj) / ((double)GR);
            const double r = bicubic_eval(c, tt, uu);
            if (v == 0) dens[p] = r;
            else if (v == 1) u0[p] = r;
            else if (v == 2) u1[p] = r;
            else u2[p] = r;
        }
    }
#pragma unroll
    for (int p = 0; p < NP; p++) vv[p] = u0[p] * u0[p] + u1[p] * u1[p] + u2[p] * u2[p];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        gather16(B.slots, n, 4 + q, p00, p10, p11, p01, xv);
        bicubic_coef(xv, c);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int ii = p / GR, jj = p % GR;
            const double tt = ((double)ii) / ((double)GR), uu = ((double)jj) / ((double)GR);
            const double fneq = bicubic_eval(c, tt, uu);
            const size_t dst = q * qs + lay0 + (size_t)(GR * jc + jj) * Xf + (GR * ic + ii);
            if constexpr (sizeof(S) == 8) {
                const double feq = feq64<q>(dens[p], u0[p], u1[p], u2[p], vv[p]);
                band[dst] = feq + rescale * fneq;
            } else {
                const double eu = edot<q>(u0[p], u1[p], u2[p]);
                const double geq = WQ[q] * ((dens[p] - 1.) + dens[p] * (3. * eu + 4.5 * eu * eu - 1.5 * vv[p]));
                store_s(band + dst, (float)(geq + rescale * fneq));
            }
        }
    });
}

template <class S>
static void launch_ctfd_prec(hipStream_t s, int GR, CtfdBand B0, CtfdBand B1, int Xc, int Yc, int nzb, double r) {
    dim3 grd((Xc * Yc + 127) / 128, 2), blk(128);
    switch (GR) {
    case 1: hipLaunchKernelGGL((k_ctfd<S, 1>), grd, blk, 0, s, B0, B1, Xc, Yc, nzb, r); break;
    case 2: hipLaunchKernelGGL((k_ctfd<S, 2>), grd, blk, 0, s, B0, B1, Xc, Yc, nzb, r); break;
    case 3: hipLaunchKernelGGL((k_ctfd<S, 3>), grd, blk, 0, s, B0, B1, Xc, Yc, nzb, r); break;
    default: hipLaunchKernelGGL((k_ctfd<S, 4>), grd, blk, 0, s, B0, B1, Xc, Yc, nzb, r); break;
    }
}

hipError_t launch_interface(hipStream_t s, dnsg_ctx *h, int iter) {
    const int Xc = h->d.Xc, Yc = h->d.Yc, n = Xc * Yc, GR = h->d.GR;
    // Hermite weights with the reference's expression order (tinterpolate_hrmt.c:20,34)
    const double tmp = (((double)iter) / ((double)GR));
    HermiteW H;
    H.h00 = (2. * tmp * tmp * tmp - 3. * tmp * tmp + 1.);
    H.h10 = (tmp * tmp * tmp - 2. * tmp * tmp + tmp);
    H.h01 = (-2. * tmp * tmp * tmp + 3. * tmp * tmp);
    H.h11 = (tmp * tmp * tmp - tmp * tmp);
    Levels L0 = {h->lev_lo[0], h->lev_lo[1], h->lev_lo[2], h->slots_lo};
    Levels L1 = {h->lev_up[0], h->lev_up[1], h->lev_up[2], h->slots_up};
    {
        dim3 grd((n * NV + 255) / 256, 2), blk(256);
        hipLaunchKernelGGL(k_hermite, grd, blk, 0, s, L0, L1, n * NV, H);
        hipLaunchKernelGGL(k_deriv12, grd, blk, 0, s, h->slots_lo, h->slots_up, Xc, Yc);
        hipLaunchKernelGGL(k_deriv3, grd, blk, 0, s, h->slots_lo, h->slots_up, Xc, Yc);
    }
    // tau_f/(GR tau_c) written with the reference's variables (omega = 1/tau):
    // tauc/((double)GR*tauf) (ctfdtransfer_opt.c:93)
    const double rescale = h->d.omega_c / ((double)GR * h->d.omega_f);
    const int zds = h->upper.nz;
    CtfdBand B0 = {h->slots_lo, h->lower.n(), zds - 1};  // lower band: interface is the top stored layer
    CtfdBand B1 = {h->slots_up, h->upper.n(), 0};        // upper band: interface is stored layer 0
    if (h->p.precision == DNSG_PREC_FP64) launch_ctfd_prec<double>(s, GR, B0, B1, Xc, Yc, zds, rescale);
    else if (h->p.precision == DNSG_PREC_FP32) launch_ctfd_prec<float>(s, GR, B0, B1, Xc, Yc, zds, rescale);
    else launch_ctfd_prec<__half>(s, GR, B0, B1, Xc, Yc, zds, rescale);
    return hipGetLastError();
}

// ---- fine -> coarse injection --------------------------------------------------------
struct FtcdSide {
    const void *band;  // band post-stream lattice
    int layer;         // stored layer coincident with the coarse end layer
    int kc;            // coarse layer written
};

template <class S>
__global__ __launch_bounds__(256) void k_ftcd(FtcdSide A0, FtcdSide A1, void *coarse, int Xc, int Yc, int zdc,
                                              int GR, int nzb, double fac) {
    const FtcdSide &A = blockIdx.y == 0 ? A0 : A1;
    const int n = Xc * Yc;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ic = id % Xc, jc = id / Xc;
    const int Xf = GR * Xc, Yf = GR * Yc;
    const size_t qf = (size_t)Xf * Yf * nzb, qc = (size_t)n * zdc;
    const S *band = static_cast<const S *>(A.band);
    S *cs = static_cast<S *>(coarse);
    const size_t src = ((size_t)A.layer * Yf + (size_t)GR * jc) * Xf + (size_t)GR * ic;
    const size_t dst = (size_t)A.kc * n + id;
    double f[Q];
    unroll<0, Q>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        if constexpr (sizeof(S) == 8) f[q] = band[q * qf + src];
        else f[q] = (double)load_s(band + q * qf + src);
    });
    if constexpr (sizeof(S) == 8) {
        const Macro64 m = macro64(f);
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const double feq = feq64<q>(m.dens, m.u0, m.u1, m.u2, m.v);
            const double fneq = f[q] - feq;
            cs[q * qc + dst] = feq + fac * fneq;
        });
    } else {
        const double drho = zeroth_moment(f), dens = 1. + drho;
        const double u0 = first_moment<0>(f) / dens, u1 = first_moment<1>(f) / dens, u2 = first_moment<2>(f) / dens;
        const double v = u0 * u0 + u1 * u1 + u2 * u2;
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const double eu = edot<q>(u0, u1, u2);
            const double geq = WQ[q] * (drho + dens * (3. * eu + 4.5 * eu * eu - 1.5 * v));
            store_s(cs + q * qc + dst, (float)(geq + fac * (f[q] - geq)));
        });
    }
}

hipError_t launch_ftcd(hipStream_t s, dnsg_ctx *h) {
    const int Xc = h->d.Xc, Yc = h->d.Yc, GR = h->d.GR, zdc = h->d.zdc, zdf = h->d.zdf;
    const int nzb = h->lower.nz;
    // lower band k = zdf-1-GR is stored layer zdf-2-GR; upper band k = GR is stored layer GR
    FtcdSide A0 = {h->lower.n(), zdf - 2 - GR, 0};
    FtcdSide A1 = {h->upper.n(), GR, zdc - 1};
    // GR*(tauf/tauc) with omega = 1/tau (ftcdtransfer.c:66)
    const double fac = GR * (h->d.omega_f / h->d.omega_c);
    dim3 grd((Xc * Yc + 255) / 256, 2), blk(256);
    void *cn = h->coarse.n();
    if (h->p.precision == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_ftcd<double>, grd, blk, 0, s, A0, A1, cn, Xc, Yc, zdc, GR, nzb, fac);
    else if (h->p.precision == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_ftcd<float>, grd, blk, 0, s, A0, A1, cn, Xc, Yc, zdc, GR, nzb, fac);
    else
        hipLaunchKernelGGL(k_ftcd<__half>, grd, blk, 0, s, A0, A1, cn, Xc, Yc, zdc, GR, nzb, fac);
    return hipGetLastError();
}

}  // namespace dnsg
