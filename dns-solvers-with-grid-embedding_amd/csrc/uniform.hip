// uniform.hip -- the '-i' start: a uniform-grid field interpolated onto the coarse
// grid and the two wall bands (initializer.c:15-545 with sp_interpolate_qd.c:22-45).
//
// The input is the ftemp half of the uniform-grid solver's lbf file, [xu][yu][zu][19]
// (zu = zdu + 2 nodes across the channel, walls at 0 and zu-1).  Every non-wall node
// of the three grids takes the triquadratic Lagrange interpolant of the 3 x 3 x 3
// uniform nodes around it (weights on the offsets -1, 0, 1, applied along z, then y,
// then x), and its non-equilibrium part is rescaled to the node's grid:
//   f = feq + (dx_u / dx_g)(omega_u / omega_g)(f_int - feq).
// One thread per destination node; x is the fastest index, so the stores of the 19
// rows are coalesced.  The uniform field's x/y halo is its periodic image.  This runs
// once per start, so the 27-point gathers go through L2 unstaged.
#include <string>

#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {

namespace {

struct UniArgs {
    const double *u;  // [xu][yu][zu][19]
    int xu, yu, zu;
    void *dst[3];     // coarse, lower band, upper band state lattices
    int Xc, Yc, zdc, Xf, Yf, zdf, GR;
    double dxf, dxc, dxr;
    double fac_c, fac_f;  // (dx_u/dx_g)(omega_u/omega_g)
};

__host__ __device__ inline int wrapm(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// reference indices (the middle of the 3) and offsets in uniform spacings of node
// (i, j, k) of grid g (0 coarse, 1 lower band, 2 upper band); the expressions keep
// initializer.c's grouping, which decides the floor()s and the last bits of the offsets
__host__ __device__ inline void uni_stencil(const UniArgs &A, int g, int i, int j, int k, int &iref, int &jref,
                                            int &k0, double &dx, double &dy, double &dz) {
    const double dxf = A.dxf, dxc = A.dxc, dxr = A.dxr;
    const int GR = A.GR, zdf = A.zdf, zdc = A.zdc;
    int kref;
    if (g == 1) {  // :326-366
        iref = (int)floor((i + 0.5) * (dxf / dxr) - 0.5);
        jref = (int)floor((j + 0.5) * (dxf / dxr) - 0.5);
        kref = (int)floor((k - 0.5) * (dxf / dxr) + 0.5);
        dx = (i + 0.5) * dxf - (iref + 0.5) * dxr;
        dy = (j + 0.5) * dxf - (jref + 0.5) * dxr;
        if (kref == 0) {  // one-sided at the wall: nodes 0, 1, 2
            k0 = 0;
            dz = -dxr + (k - 0.5) * dxf + 0.5 * dxr - kref * dxr;
        } else {
            k0 = kref - 1;
            dz = (k - 0.5) * dxf + 0.5 * dxr - kref * dxr;
        }
    } else if (g == 0) {  // :403-431
        iref = (int)floor((i + 0.5) * dxc / dxr - 0.5);
        jref = (int)floor((j + 0.5) * dxc / dxr - 0.5);
        kref = (int)floor(0.5 + (k * GR + (zdf - GR - 1 - 0.5)) * dxf / dxr);
        dx = (i + 0.5) * dxc - (iref + 0.5) * dxr;
        dy = (j + 0.5) * dxc - (jref + 0.5) * dxr;
        k0 = kref - 1;
        dz = (k * GR + (zdf - GR - 1 - 0.5)) * dxf - (kref - 0.5) * dxr;
    } else {  // :470-499
        iref = (int)floor((i + 0.5) * dxf / dxr - 0.5);
        jref = (int)floor((j + 0.5) * dxf / dxr - 0.5);
        kref = (int)floor(0.5 + (k + GR * (zdc - 3) + zdf - 1 - 0.5) * dxf / dxr);
        dx = (i + 0.5) * dxf - (iref + 0.5) * dxr;
        dy = (j + 0.5) * dxf - (jref + 0.5) * dxr;
        k0 = kref - 1;
        dz = (k + GR * (zdc - 3) + zdf - 1 - 0.5) * dxf - (kref - 0.5) * dxr;
    }
    dx = dx / dxr;
    dy = dy / dxr;
    dz = dz / dxr;
}

// destination node ranges: lower band k = 1 .. zdf-1 (stored layer k-1), upper band
// k = 0 .. zdf-2 (stored layer k), coarse k = 0 .. zdc-1
__host__ __device__ inline void uni_grid(const UniArgs &A, int g, int &nx, int &ny, int &nk, int &klo) {
    nx = g == 0 ? A.Xc : A.Xf;
    ny = g == 0 ? A.Yc : A.Yf;
    nk = g == 0 ? A.zdc : A.zdf - 1;
    klo = g == 1 ? 1 : 0;
}

template <class S>
__global__ __launch_bounds__(256) void k_uniform(UniArgs A) {
    const size_t nc = (size_t)A.Xc * A.Yc * A.zdc, nb = (size_t)A.Xf * A.Yf * (A.zdf - 1);
    size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= nc + 2 * nb) return;
    int g = 0;
    if (id >= nc) {
        id -= nc;
        g = 1 + (id >= nb ? 1 : 0);
        if (id >= nb) id -= nb;
    }
    int nx, ny, nk, klo;
    uni_grid(A, g, nx, ny, nk, klo);
    const int i = (int)(id % nx), j = (int)((id / nx) % ny), m = (int)(id / ((size_t)nx * ny));
    const int k = m + klo;
    int iref, jref, k0;
    double dx, dy, dz;
    uni_stencil(A, g, i, j, k, iref, jref, k0, dx, dy, dz);
    // sp_interpolate_qd.c:27-44: the Lagrange weights, grouped as there
    const double wz0 = 0.5 * dz * (dz - 1.), wz1 = (dz - 1.) * (dz + 1.), wz2 = 0.5 * (dz + 1.) * dz;
    const double wy0 = 0.5 * dy * (dy - 1.), wy1 = (dy - 1.) * (dy + 1.), wy2 = 0.5 * (dy + 1.) * dy;
    const double wx0 = 0.5 * dx * (dx - 1.), wx1 = (dx - 1.) * (dx + 1.), wx2 = 0.5 * (dx + 1.) * dx;
    const double *col[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
            col[a][b] = A.u + (((size_t)wrapm(iref - 1 + a, A.xu) * A.yu + wrapm(jref - 1 + b, A.yu)) * A.zu + k0) * Q;
    double fi[Q];
    unroll<0, Q>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        double cy[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double cz[3];
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const double *p = col[a][b] + q;
                cz[b] = wz0 * p[0] - wz1 * p[Q] + wz2 * p[2 * Q];
            }
            cy[a] = wy0 * cz[0] - wy1 * cz[1] + wy2 * cz[2];
        }
        fi[q] = wx0 * cy[0] - wx1 * cy[1] + wx2 * cy[2];
    });
    // :367-392: rescale the non-equilibrium part to the destination grid
    const Macro64 mc = macro64(fi);
    const double fac = g == 0 ? A.fac_c : A.fac_f;
    S *dst = static_cast<S *>(A.dst[g]);
    const size_t row = (size_t)m * ny + j;
    unroll<0, Q>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        const double feq = feq64<q>(mc.dens, mc.u0, mc.u1, mc.u2, mc.v);
        const double out = feq + fac * (fi[q] - feq);
        if constexpr (sizeof(S) == 8) dst[lat_at(q, row, i, nx)] = out;
        else store_s(dst + lat_at(q, row, i, nx), (float)(out - WQ[q]));
    });
}

UniArgs uni_args(const dnsg_ctx *h, const double *u, int xdu, int ydu, int zdu) {
    const dnsg_dims &d = h->d;
    UniArgs A{};
    A.u = u;
    A.xu = xdu;
    A.yu = ydu;
    A.zu = zdu + 2;  // main:140
    A.dst[0] = h->coarse.c();
    A.dst[1] = h->lower.c();
    A.dst[2] = h->upper.c();
    A.Xc = d.Xc; A.Yc = d.Yc; A.zdc = d.zdc; A.Xf = d.Xf; A.Yf = d.Yf; A.zdf = d.zdf; A.GR = d.GR;
    A.dxf = 2. / ((double)(d.NZs - 2.));  // initializer.c:49-51
    A.dxc = d.GR * A.dxf;
    A.dxr = 2. / ((double)(A.zu - 2.));
    // main:443-444: the uniform grid's omega from the same Reynolds number and bulk velocity
    const double omega_u = 1. / ((3. / h->p.Rebs) * h->p.ub * (A.zu - 2.) + 0.5);
    A.fac_c = (A.dxr / A.dxc) * (omega_u / d.omega_c);
    A.fac_f = (A.dxr / A.dxf) * (omega_u / d.omega_f);
    return A;
}

}  // namespace

// every 3 x 3 x 3 stencil must lie inside the uniform array and the halo the
// reference allocates around it (x, y in [-2, n+3], z in [0, zu-1]); the indices
// are monotone in i, j, k, so the faces of each grid's node box bound them
int uniform_check(const dnsg_ctx *h, int xdu, int ydu, int zdu, std::string *err) {
    if (xdu < 1 || ydu < 1 || zdu < 2) {
        *err = "uniform grid sizes must be xdu >= 1, ydu >= 1, zdu >= 2";
        return DNSG_ERR_ARG;
    }
    const UniArgs A = uni_args(h, nullptr, xdu, ydu, zdu);
    for (int g = 0; g < 3; g++) {
        int nx, ny, nk, klo;
        uni_grid(A, g, nx, ny, nk, klo);
        const int ic[2] = {0, nx - 1}, jc[2] = {0, ny - 1}, kc[2] = {klo, klo + nk - 1};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++) {
                    int iref, jref, k0;
                    double dx, dy, dz;
                    uni_stencil(A, g, ic[a], jc[b], kc[c], iref, jref, k0, dx, dy, dz);
                    if (iref - 1 < -2 || iref + 1 > xdu + 3 || jref - 1 < -2 || jref + 1 > ydu + 3 || k0 < 0 ||
                        k0 + 2 > A.zu - 1) {
                        *err = "the uniform grid " + std::to_string(xdu) + "x" + std::to_string(ydu) + "x" +
                               std::to_string(A.zu) + " does not cover grid " + std::to_string(g) +
                               " (its interpolation stencils leave the uniform array)";
                        return DNSG_ERR_ARG;
                    }
                }
    }
    return DNSG_OK;
}

hipError_t launch_uniform(hipStream_t s, dnsg_ctx *h, const double *u, int xdu, int ydu, int zdu) {
    const UniArgs A = uni_args(h, u, xdu, ydu, zdu);
    const size_t n = (size_t)A.Xc * A.Yc * A.zdc + 2 * (size_t)A.Xf * A.Yf * (A.zdf - 1);
    const dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    if (h->p.precision == DNSG_PREC_FP64) hipLaunchKernelGGL(k_uniform<double>, grd, blk, 0, s, A);
    else if (h->p.precision == DNSG_PREC_FP32) hipLaunchKernelGGL(k_uniform<float>, grd, blk, 0, s, A);
    else hipLaunchKernelGGL(k_uniform<__half>, grd, blk, 0, s, A);
    return hipGetLastError();
}

}  // namespace dnsg
