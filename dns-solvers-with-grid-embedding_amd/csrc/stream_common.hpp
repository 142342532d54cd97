// stream_common.hpp -- device helpers shared by the collide-stream kernels
// (k_stream.hip) and the two-substep band kernel (k_band2.hip): the compute type
// of each storage type, the collision dispatch, the fused-reduction sums and the
// push of one collided node to its 19 targets with the wall rules.
#pragma once
#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {

template <class S>
struct ComputeOf {
    using T = float;
};
template <>
struct ComputeOf<double> {
    using T = double;
};

__device__ __forceinline__ void collide_any(double *f, double G, double om) { collide64(f, G, om); }
__device__ __forceinline__ void collide_any(float *f, float G, float om) { collide32(f, G, om); }

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap the linear block
// id so that each XCD gets one contiguous range of (row, chunk) pairs.  The
// neighbouring chunks of a row, whose e_x-shifted writes share cache lines,
// then meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int per = nb >> 3;
    return (b < (per << 3)) ? (b & 7) * per + (b >> 3) : b;
}

// fused-reduction sums of one thread's nodes: pre-collision moments (rho, mom, u)
// and the destination-weighted sums of what it writes (dm, dd)
struct RowSums {
    double rho = 0., mom = 0., u = 0., dm = 0., dd = 0.;
    template <class S, class T>
    __device__ __forceinline__ void add(const T *f) {
        double d, m;
        if constexpr (sizeof(S) == 8) {
            m = first_moment<0>(f);
            d = zeroth_moment(f);
        } else {
            m = (double)first_moment<0>(f);
            d = 1. + (double)zeroth_moment(f);
        }
        rho += d;
        mom += m;
        u += m / d;
    }
};

// encalc2 weight of a destination layer (LatArgs::dsum)
__device__ __forceinline__ double dest_w(const LatArgs &L, int zt) {
    return (zt < L.klo || zt > L.khi) ? 0. : (zt == L.khalf ? 0.5 * L.kscale : L.kscale);
}

// add a store of direction d (e_x = ex) and f-value fv into a layer of weight W
__device__ __forceinline__ void add_write(RowSums &rs, double W, int ex, double fv) {
    const double a = W * fv;
    rs.dd += a;
    if (ex > 0) rs.dm += a;
    else if (ex < 0) rs.dm -= a;
}


// Push of the collided node at x of row (y, z) = row: each direction to its target
// (x + e_x, y + e_y, z + e_z) with the boundary rules at the stored z ends
// (xfcomputations.c:215-442):
//   BOUND_DROP  populations leaving the lattice are not written, nor are those
//               entering a skipped layer (rewritten by an interface transfer)
//   BOUND_WALL  target on the wall with code s(t) in {-1, -2}, s1 the code below
//               the source: s(t) = -1 and s1 = -1 -> halfway bounce-back
//               f(x, opp q); otherwise specular reflection f(x + e_xy, mirror q)
// With DS and L.dsum, the destination-weighted sums of the writes go into rs.
// The wall codes come from wcode(zs, dx, dy, xt, yt): the code of the wall at the
// in-plane offset (dx, dy) from the node, (xt, yt) in lattice coordinates, on the lo
// (zs < 0) or hi wall.
struct WallCodesGlobal {
    const LatArgs &L;
    __device__ __forceinline__ int operator()(int zs, int, int, int xt, int yt) const {
        const int8_t *wc = zs < 0 ? L.wlo : L.whi;
        return wc[(size_t)yt * L.nx + xt];
    }
};

template <class S, bool DS, class T, class WC>
__device__ __forceinline__ void scatter_push(const LatArgs &L, S *__restrict__ nxt, int row, int x, const T *f,
                                             RowSums &rs, const WC &wcode) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int y = row % ny, z = row / ny;
    const bool lo_edge = (z == 0), hi_edge = (z == nz - 1);
    const int ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const bool ds = DS && L.dsum != nullptr;  // DS: compiled in only where a launch asks for it
    const double wz[3] = {ds ? dest_w(L, z - 1) : 0., ds ? dest_w(L, z) : 0., ds ? dest_w(L, z + 1) : 0.};
    const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int xt = (EX[q] > 0) ? xp : (EX[q] < 0 ? xm : x);
        const int yt = (EY[q] > 0) ? yp : (EY[q] < 0 ? ym : y);
        if constexpr (EZ[q] != 0) {
            const bool at_end = (EZ[q] < 0) ? lo_edge : hi_edge;
            if (at_end) {
                const int kind = (EZ[q] < 0) ? L.lo_kind : L.hi_kind;
                if (kind == BOUND_WALL) {
                    const int sld = wcode(EZ[q], EX[q], EY[q], xt, yt);
                    const int sld1 = wcode(EZ[q], 0, 0, x, y);
                    if (sld == -1 && sld1 == -1) {
                        store_s(nxt + lat_at(OPP[q], (size_t)row, x, nx), f[q]);
                        if (ds) add_write(rs, wz[1], -EX[q], read_back<S, q>(f[q]));
                    } else {
                        store_s(nxt + lat_at(MIR[q], (size_t)z * ny + yt, xt, nx), f[q]);
                        if (ds) add_write(rs, wz[1], EX[q], read_back<S, q>(f[q]));
                    }
                }
                return;
            }
        }
        const int zt = z + EZ[q];
        if ((zt == 0 && L.skip_lo) || (zt == nz - 1 && L.skip_hi)) return;
        store_s(nxt + lat_at(q, (size_t)zt * ny + yt, xt, nx), f[q]);
        if (ds) add_write(rs, wz[EZ[q] + 1], EX[q], read_back<S, q>(f[q]));
    });
}

}  // namespace dnsg
