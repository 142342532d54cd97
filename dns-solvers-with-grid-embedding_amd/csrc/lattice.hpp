// lattice.hpp -- D3Q19 tables and the BGK collision, written for CDNA (gfx950).
//
// Direction numbering, weights and the forcing weight BI are the reference's
// (globals.c:3-13).  The FP64 collision reproduces the reference's expression
// order (xcomputations.c:59-85) operation by operation; terms the reference
// multiplies by a zero lattice component are skipped, which only changes the
// sign of an exact zero in an intermediate and never a result.  The build uses
// -ffp-contract=off, so no multiply-add is fused and every product rounds as on
// the host.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace dnsg {

constexpr int Q = 19;
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};
constexpr int MIR[Q] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 12, 11, 14, 13, 16, 15, 18, 17};
// 1/3, 1/18, 1/36 and 3*w, rounded once from the exact fractions as the
// reference's initialisers are
constexpr double W0 = 1. / 3., W1 = 1. / 18., W2 = 1. / 36.;
constexpr double WQ[Q] = {W0, W1, W1, W1, W1, W1, W1, W2, W2, W2, W2, W2, W2, W2, W2, W2, W2, W2, W2};
constexpr double BIQ[Q] = {0., 1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 6., 1. / 12., 1. / 12.,
                           1. / 12., 1. / 12., 1. / 12., 1. / 12., 1. / 12., 1. / 12., 1. / 12., 1. / 12.,
                           1. / 12., 1. / 12.};

// compile-time loop
template <int I, int N, class F>
__device__ __forceinline__ void unroll(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        unroll<I + 1, N>(f);
    }
}

// sum_a E_c[a] * f[a] in the reference's order a = 0..18, starting from 0
template <int C, class T>
__device__ __forceinline__ T first_moment(const T *f) {
    T s = T(0);
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int e = (C == 0) ? EX[q] : (C == 1) ? EY[q] : EZ[q];
        if constexpr (e > 0) s = s + f[q];
        else if constexpr (e < 0) s = s - f[q];
    });
    return s;
}

template <class T>
__device__ __forceinline__ T zeroth_moment(const T *f) {
    T s = T(0);
    unroll<0, Q>([&](auto qc) { s = s + f[decltype(qc)::value]; });
    return s;
}

// e_q . u as (E0*u0 + E1*u1) + E2*u2 with the zero terms dropped
template <int q, class T>
__device__ __forceinline__ T edot(T u0, T u1, T u2) {
    T s = T(0);
    bool have = false;
    auto add = [&](int e, T u) {
        if (e == 0) return;
        T t = (e > 0) ? u : -u;
        s = have ? s + t : t;
        have = true;
    };
    add(EX[q], u0);
    add(EY[q], u1);
    add(EZ[q], u2);
    return s;
}

// feq = w*rho*(1 + 3eu + 4.5eu^2 - 1.5v), grouped as the reference groups it
template <int q>
__device__ __forceinline__ double feq64(double dens, double u0, double u1, double u2, double v) {
    const double eu = edot<q>(u0, u1, u2);
    return WQ[q] * dens * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * v);
}

// Macroscopic moments of one node: rho, u = (sum e f)/rho, v = |u|^2.
struct Macro64 {
    double dens, u0, u1, u2, v;
};

__device__ __forceinline__ Macro64 macro64(const double *f) {
    Macro64 m;
    const double ux = first_moment<0>(f), uy = first_moment<1>(f), uz = first_moment<2>(f);
    m.dens = zeroth_moment(f);
    m.u0 = ux / m.dens;
    m.u1 = uy / m.dens;
    m.u2 = uz / m.dens;
    m.v = m.u0 * m.u0 + m.u1 * m.u1 + m.u2 * m.u2;
    return m;
}

// BGK collision with the first-order body force F = 3 w rho e_x G
// (xcomputations.c:73-85): f <- f + (F - omega (f - feq)).
__device__ __forceinline__ void collide64(double *f, double G, double omega) {
    const Macro64 m = macro64(f);
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double feq = feq64<q>(m.dens, m.u0, m.u1, m.u2, m.v);
        const double t = omega * (f[q] - feq);
        if constexpr (EX[q] == 0) {
            f[q] = f[q] - t;
        } else {
            const double Fi = BIQ[q] * m.dens * (EX[q] > 0 ? G : -G);
            f[q] = f[q] + (Fi - t);
        }
    });
}

// ---- precision tier: shifted storage g = f - w*rho0, FP32 collision ---------
// With rho0 = 1: rho = 1 + sum g, rho u = sum e g (sum_q w_q e_q = 0), and
// geq = feq - w = w*(drho + rho*(3eu + 4.5eu^2 - 1.5v)), which keeps the
// O(1) part out of the FP32 arithmetic.
struct Macro32 {
    float drho, dens, u0, u1, u2, v;
};

__device__ __forceinline__ Macro32 macro32(const float *g) {
    Macro32 m;
    m.drho = zeroth_moment(g);
    m.dens = 1.f + m.drho;
    const float inv = 1.f / m.dens;
    m.u0 = first_moment<0>(g) * inv;
    m.u1 = first_moment<1>(g) * inv;
    m.u2 = first_moment<2>(g) * inv;
    m.v = m.u0 * m.u0 + m.u1 * m.u1 + m.u2 * m.u2;
    return m;
}

template <int q>
__device__ __forceinline__ float geq32(const Macro32 &m) {
    const float eu = edot<q>(m.u0, m.u1, m.u2);
    return float(WQ[q]) * (m.drho + m.dens * (3.f * eu + 4.5f * eu * eu - 1.5f * m.v));
}

__device__ __forceinline__ void collide32(float *g, float G, float omega) {
    const Macro32 m = macro32(g);
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float t = omega * (g[q] - geq32<q>(m));
        if constexpr (EX[q] == 0) {
            g[q] = g[q] - t;
        } else {
            const float Fi = float(BIQ[q]) * m.dens * (EX[q] > 0 ? G : -G);
            g[q] = g[q] + (Fi - t);
        }
    });
}

// storage conversions for the shifted tiers
__device__ __forceinline__ float load_s(const float *p) { return *p; }
__device__ __forceinline__ float load_s(const __half *p) { return __half2float(*p); }
__device__ __forceinline__ double load_s(const double *p) { return *p; }
__device__ __forceinline__ void store_s(float *p, float v) { *p = v; }
__device__ __forceinline__ void store_s(__half *p, float v) { *p = __float2half(v); }
__device__ __forceinline__ void store_s(double *p, double v) { *p = v; }
// the value a later read of a store returns, as f in FP64 (the shifted tiers
// store g = f - w, FP16 after rounding to half)
template <class S, int q, class T>
__device__ __forceinline__ double read_back(T v) {
    if constexpr (sizeof(S) == 8) return (double)v;
    else if constexpr (sizeof(S) == 4) return (double)(float)v + WQ[q];
    else return (double)__half2float(__float2half((float)v)) + WQ[q];
}
// streaming stores that bypass the caches' normal allocation (written once, read next step)
__device__ __forceinline__ void store_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(__half *p, float v) {
    const __half h = __float2half(v);
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned short *>(&h), reinterpret_cast<unsigned short *>(p));
}

}  // namespace dnsg
