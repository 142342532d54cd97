// k_stream.hip -- fused BGK collide + push-stream on the coarse grid and the
// wall bands, the initial-state kernel and the AoS <-> SoA transposes.
//
// One workgroup owns a chunk of CS_THREADS nodes of one (y, z) row: its reads
// are 19 contiguous vectors of the row block, its writes the same vectors of
// the 9 neighbouring rows shifted by e_x.  A lattice narrower than CS_THREADS
// packs CS_THREADS / nx whole rows into each workgroup instead.
//
// Boundary rules at the stored z ends (xfcomputations.c:215-442):
//   BOUND_DROP  populations leaving the lattice are not written (the bounded
//               layers of the reference: coarse k = 0, zdc-1 and the band
//               interface layers, which the interface transfer rewrites)
//   BOUND_WALL  the target is a wall node with code s(t) in {-1, -2}; with
//               s1 the code below the source node:
//                 s(t) = -1 and s1 = -1  -> halfway bounce-back  f(x, opp q)
//                 otherwise              -> specular reflection  f(x + e_xy, mirror q)
#include "internal.hpp"
#include "lattice.hpp"
#include "stream_common.hpp"

namespace dnsg {

// Per storage type: use the aligned-store form (bit 0 FP64, 1 FP32, 2 FP16) and,
// in it, non-temporal stores.  Measured on the Re_tau=180 lattices: the aligned
// form with non-temporal stores is fastest in FP64; FP32 and FP16 are fastest in
// the scatter form.
#ifndef DNSG_CS_ALIGNED_MASK
#define DNSG_CS_ALIGNED_MASK 1
#endif
#ifndef DNSG_CS_NT_MASK
#define DNSG_CS_NT_MASK 1
#endif
template <class S>
constexpr int prec_bit() { return sizeof(S) == 8 ? 1 : (sizeof(S) == 4 ? 2 : 4); }
template <class S>
constexpr bool cs_aligned() { return (DNSG_CS_ALIGNED_MASK & prec_bit<S>()) != 0; }
template <class S>
constexpr bool cs_nt() { return (DNSG_CS_NT_MASK & prec_bit<S>()) != 0; }
template <class S, class V>
__device__ __forceinline__ void store_cs(S *p, V v) {
    if constexpr (cs_nt<S>()) store_nt(p, v);
    else store_s(p, v);
}
// Non-temporal loads of the pre-collision lattice (same mask convention) measured
// 1.7% slower in FP64 and no faster in FP32 on the Re_tau=180 lattices: off.
#ifndef DNSG_CS_NTLOAD_MASK
#define DNSG_CS_NTLOAD_MASK 0
#endif
template <class S>
__device__ __forceinline__ auto load_cs(const S *p) {
    if constexpr ((DNSG_CS_NTLOAD_MASK & prec_bit<S>()) != 0 && sizeof(S) != 2) return __builtin_nontemporal_load(p);
    else return load_s(p);
}

// Scatter form for the node at x of row (y, z): collide, then write each
// direction to its target (x + e_x, y + e_y, z + e_z) with the wall rules.
template <class S, bool DS, class T>
__device__ __forceinline__ void cs_scatter(const LatArgs &L, const S *__restrict__ cur, S *__restrict__ nxt, int row,
                                           int x, T G, T om, bool red, RowSums &rs) {
    const int nx = L.nx;
    T f[Q];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = load_cs(cur + (size_t)q * nx + x);
    });
    if (red) rs.add<S>(f);
    collide_any(f, G, om);
    scatter_push<S, DS>(L, nxt, row, x, f, rs, WallCodesGlobal{L});
}

// Aligned form for chunk x0 .. x0 + CS_THREADS - 1 of row (y, z) (nx a multiple of
// CS_THREADS): the workgroup also collides the two nodes just outside its chunk
// and passes the e_x != 0 values one lane along (shuffles within a wave, LDS
// between waves), so that every store of direction q lands on the workgroup's own
// positions of the target row: whole cache lines, written once.  Each thread
// stores, at its own x, the value arriving there from x - e_x; the set of writes
// is exactly the scatter form's.
template <class S, bool DS, class T>
__device__ __forceinline__ void cs_aligned_chunk(const LatArgs &L, const S *__restrict__ cur, S *__restrict__ nxt,
                                                 int row, int x0, T G, T om, bool red, RowSums &rs) {
    constexpr int NW = CS_THREADS / 64;
    __shared__ T hb[2][Q];      // post-collision halo nodes x0 - 1, x0 + CS_THREADS
    __shared__ T wb[NW][2][Q];  // per wave: last lane's e_x > 0, first lane's e_x < 0 values
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int y = row % ny, z = row / ny;
    const bool lo_edge = (z == 0), hi_edge = (z == nz - 1);
    const int ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, x = x0 + t;
    if (t == 0 || t == CS_THREADS - 1) {
        const int side = t == 0 ? 0 : 1;
        int xh = side == 0 ? x0 - 1 : x0 + CS_THREADS;
        xh = xh < 0 ? xh + nx : (xh >= nx ? xh - nx : xh);
        T g[Q];
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            g[q] = load_cs(cur + (size_t)q * nx + xh);
        });
        collide_any(g, G, om);
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (EX[q] != 0) hb[side][q] = g[q];
        });
    }
    T f[Q];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = load_cs(cur + (size_t)q * nx + x);
    });
    if (red) rs.add<S>(f);
    collide_any(f, G, om);
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (EX[q] > 0) {
            if (lane == 63) wb[w][0][q] = f[q];
        } else if constexpr (EX[q] < 0) {
            if (lane == 0) wb[w][1][q] = f[q];
        }
    });
    __syncthreads();
    const bool ds = DS && L.dsum != nullptr;  // DS: compiled in only where a launch asks for it
    const double wz[3] = {ds ? dest_w(L, z - 1) : 0., ds ? dest_w(L, z) : 0., ds ? dest_w(L, z + 1) : 0.};
    const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        T v = f[q];  // the value arriving at x: post-collision f_q of node x - e_x
        if constexpr (EX[q] > 0) {
            v = __shfl_up(f[q], 1, 64);
            if (lane == 0) v = (w == 0) ? hb[0][q] : wb[w - 1][0][q];
        } else if constexpr (EX[q] < 0) {
            v = __shfl_down(f[q], 1, 64);
            if (lane == 63) v = (w == NW - 1) ? hb[1][q] : wb[w + 1][1][q];
        }
        const int yt = (EY[q] > 0) ? yp : (EY[q] < 0 ? ym : y);
        if constexpr (EZ[q] != 0) {
            const bool at_end = (EZ[q] < 0) ? lo_edge : hi_edge;
            if (at_end) {
                const int kind = (EZ[q] < 0) ? L.lo_kind : L.hi_kind;
                if (kind == BOUND_WALL) {
                    const int8_t *wc = (EZ[q] < 0) ? L.wlo : L.whi;
                    // specular reflection of the source x - e_x into (x, yt), unless it bounces back
                    const int xs = (EX[q] > 0) ? xm : (EX[q] < 0 ? xp : x);
                    if (!(wc[(size_t)yt * nx + x] == -1 && wc[(size_t)y * nx + xs] == -1)) {
                        store_cs(nxt + lat_at(MIR[q], (size_t)z * ny + yt, x, nx), v);
                        if (ds) add_write(rs, wz[1], EX[q], read_back<S, q>(v));
                    }
                    // halfway bounce-back of this node's own value
                    const int xt = (EX[q] > 0) ? xp : (EX[q] < 0 ? xm : x);
                    if (wc[(size_t)yt * nx + xt] == -1 && wc[(size_t)y * nx + x] == -1) {
                        store_cs(nxt + lat_at(OPP[q], (size_t)row, x, nx), f[q]);
                        if (ds) add_write(rs, wz[1], -EX[q], read_back<S, q>(f[q]));
                    }
                }
                return;
            }
        }
        const int zt = z + EZ[q];
        if ((zt == 0 && L.skip_lo) || (zt == nz - 1 && L.skip_hi)) return;
        store_cs(nxt + lat_at(q, (size_t)zt * ny + yt, x, nx), v);
        if (ds) add_write(rs, wz[EZ[q] + 1], EX[q], read_back<S, q>(v));
    });
}

// Paired form (FP32 / FP16 storage, nx a multiple of 2 CS_THREADS): the workgroup
// takes nodes x0 .. x0 + 2 CS_THREADS - 1, thread t the pair x = x0 + 2t, x + 1, read
// and written as one 2-value access per direction.  As in the aligned form, the
// workgroup also collides the nodes x0 - 1 and x0 + 2 CS_THREADS, and each thread
// stores at (x, x + 1) the values arriving there from (x, x + 1) - e_x: for e_x = +1
// the pair (f(x - 1), f(x)), f(x - 1) being the previous thread's second node; for
// e_x = -1 the pair (f(x + 1), f(x + 2)).  At a wall end the pair's two positions
// take the wall rules one value at a time.  The writes are exactly the scatter
// form's.  rs[0] / rs[1] collect the sums of the first / second chunk's nodes.
__device__ __forceinline__ void load_pair(const float *p, float &a, float &b) {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void load_pair(const __half *p, float &a, float &b) {
    const __half2 v = *reinterpret_cast<const __half2 *>(p);
    a = __low2float(v);
    b = __high2float(v);
}
__device__ __forceinline__ void load_pair(const double *p, double &a, double &b) {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    a = v.x;
    b = v.y;
}
// Paired stores are non-temporal, like the aligned form's.  Measured on the
// Re_tau=180 lattices: the FP64 coarse launch of the flux-controlled step 1.904 ->
// 1.861 ms; FP32 coarse 0.952 -> 0.938 ms and band 0.479 -> 0.472 ms; FP16 within 1%.
__device__ __forceinline__ void store_pair2(double *p, double a, double b) {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p));
}
__device__ __forceinline__ void store_pair2(float *p, float a, float b) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<f2v *>(p));
}
__device__ __forceinline__ void store_pair2(__half *p, float a, float b) {
    const __half2 h = __halves2half2(__float2half(a), __float2half(b));
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned int *>(&h), reinterpret_cast<unsigned int *>(p));
}

template <class S, bool DS, class T>
__device__ __forceinline__ void cs_paired_chunk(const LatArgs &L, const S *__restrict__ cur, S *__restrict__ nxt,
                                                int row, int x0, T G, T om, bool red, RowSums *rs) {
    constexpr int NW = CS_THREADS / 64, NP = 2 * CS_THREADS;
    __shared__ T hb[2][Q];      // post-collision nodes x0 - 1, x0 + NP
    __shared__ T wb[NW][2][Q];  // per wave: last lane's second node e_x > 0, first lane's first node e_x < 0
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int y = row % ny, z = row / ny;
    const bool lo_edge = (z == 0), hi_edge = (z == nz - 1);
    const int ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, x = x0 + 2 * t;
    if (t == 0 || t == CS_THREADS - 1) {
        const int side = t == 0 ? 0 : 1;
        int xh = side == 0 ? x0 - 1 : x0 + NP;
        xh = xh < 0 ? xh + nx : (xh >= nx ? xh - nx : xh);
        T g[Q];
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            g[q] = load_s(cur + (size_t)q * nx + xh);
        });
        collide_any(g, G, om);
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (EX[q] != 0) hb[side][q] = g[q];
        });
    }
    T f0[Q], f1[Q];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        load_pair(cur + (size_t)q * nx + x, f0[q], f1[q]);
    });
    RowSums &rt = rs[t < CS_THREADS / 2 ? 0 : 1];
    if (red) {
        rt.add<S>(f0);
        rt.add<S>(f1);
    }
    collide_any(f0, G, om);
    collide_any(f1, G, om);
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (EX[q] > 0) {
            if (lane == 63) wb[w][0][q] = f1[q];
        } else if constexpr (EX[q] < 0) {
            if (lane == 0) wb[w][1][q] = f0[q];
        }
    });
    __syncthreads();
    const bool ds = DS && L.dsum != nullptr;  // DS: compiled in only where a launch asks for it
    const double wz[3] = {ds ? dest_w(L, z - 1) : 0., ds ? dest_w(L, z) : 0., ds ? dest_w(L, z + 1) : 0.};
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        // the values arriving at x and x + 1, and the sources they leave from
        T va = f0[q], vb = f1[q];
        if constexpr (EX[q] > 0) {
            T s = __shfl_up(f1[q], 1, 64);
            if (lane == 0) s = (w == 0) ? hb[0][q] : wb[w - 1][0][q];
            va = s;
            vb = f0[q];
        } else if constexpr (EX[q] < 0) {
            T s = __shfl_down(f0[q], 1, 64);
            if (lane == 63) s = (w == NW - 1) ? hb[1][q] : wb[w + 1][1][q];
            va = f1[q];
            vb = s;
        }
        const int yt = (EY[q] > 0) ? yp : (EY[q] < 0 ? ym : y);
        if constexpr (EZ[q] != 0) {
            const bool at_end = (EZ[q] < 0) ? lo_edge : hi_edge;
            if (at_end) {
                const int kind = (EZ[q] < 0) ? L.lo_kind : L.hi_kind;
                if (kind == BOUND_WALL) {
                    const int8_t *wc = (EZ[q] < 0) ? L.wlo : L.whi;
                    // positions x + i (i = 0, 1): the arriving value (source x + i - e_x)
                    // is reflected specularly into (x + i, yt) unless it bounces back,
                    // and the node's own value bounces back where its target is no-slip
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        const int xi = x + i;
                        const int xm = (xi == 0) ? nx - 1 : xi - 1, xp = (xi == nx - 1) ? 0 : xi + 1;
                        const T v = i == 0 ? va : vb, own = i == 0 ? f0[q] : f1[q];
                        const int xs = (EX[q] > 0) ? xm : (EX[q] < 0 ? xp : xi);
                        if (!(wc[(size_t)yt * nx + xi] == -1 && wc[(size_t)y * nx + xs] == -1)) {
                            store_s(nxt + lat_at(MIR[q], (size_t)z * ny + yt, xi, nx), v);
                            if (ds) add_write(rt, wz[1], EX[q], read_back<S, q>(v));
                        }
                        const int xt = (EX[q] > 0) ? xp : (EX[q] < 0 ? xm : xi);
                        if (wc[(size_t)yt * nx + xt] == -1 && wc[(size_t)y * nx + xi] == -1) {
                            store_s(nxt + lat_at(OPP[q], (size_t)row, xi, nx), own);
                            if (ds) add_write(rt, wz[1], -EX[q], read_back<S, q>(own));
                        }
                    }
                }
                return;
            }
        }
        const int zt = z + EZ[q];
        if ((zt == 0 && L.skip_lo) || (zt == nz - 1 && L.skip_hi)) return;
        store_pair2(nxt + lat_at(q, (size_t)zt * ny + yt, x, nx), va, vb);
        if (ds) {
            add_write(rt, wz[EZ[q] + 1], EX[q], read_back<S, q>(va));
            add_write(rt, wz[EZ[q] + 1], EX[q], read_back<S, q>(vb));
        }
    });
}

__device__ __forceinline__ void write_sums(double *red, size_t e, const RowSums &s) {
    red[e + 0] = s.rho;
    red[e + 1] = s.mom;
    red[e + 2] = s.u;
}

// workgroup sum of one partial (wave shuffles, then LDS over the waves)
__device__ __forceinline__ RowSums block_sums(RowSums s) {
    __shared__ double lds[5][CS_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    s.rho = wave_sum64(s.rho);
    s.mom = wave_sum64(s.mom);
    s.u = wave_sum64(s.u);
    s.dm = wave_sum64(s.dm);
    s.dd = wave_sum64(s.dd);
    if (lane == 0) {
        lds[0][wid] = s.rho; lds[1][wid] = s.mom; lds[2][wid] = s.u; lds[3][wid] = s.dm; lds[4][wid] = s.dd;
    }
    __syncthreads();
    RowSums t;
    if (threadIdx.x == 0)
        for (int w = 0; w < CS_THREADS / 64; w++) {
            t.rho += lds[0][w]; t.mom += lds[1][w]; t.u += lds[2][w]; t.dm += lds[3][w]; t.dd += lds[4][w];
        }
    return t;
}

template <class S, bool DS>
__global__ __launch_bounds__(CS_THREADS) void k_collide_stream(Batch<LatArgs> P) {
    using T = typename ComputeOf<S>::T;
    if (halted(P.halt)) return;
    const int bg = xcd_remap(blockIdx.x, gridDim.x);
    const int it = batch_find(P.first, P.n, bg);
    const LatArgs &L = P.it[it];
    const int nx = L.nx, ny = L.ny;
    const int b = bg - P.first[it];
    const S *__restrict__ base = static_cast<const S *>(L.cur);
    S *__restrict__ nxt = static_cast<S *>(L.nxt);
    const T G = (T)(L.gscale * L.force[0]);
    const T om = (T)L.omega;
    const bool red = L.red != nullptr;
    const int rpw = cs_rows_per_wg(nx);
    RowSums rs;
    {
        if (L.paired) {
            // two chunks of one row per workgroup, two nodes per thread
            const int cpr = cs_chunks(nx), ppr = cpr / 2;
            const int row = b / ppr, pc = b % ppr;
            RowSums r2[2];
            cs_paired_chunk<S, DS, T>(L, base + (size_t)row * Q * nx, nxt, row, pc * 2 * CS_THREADS, G, om, red, r2);
            if (red || (DS && L.dsum)) {  // uniform per workgroup
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (c) __syncthreads();  // block_sums' shared partials are reused
                    const RowSums t = block_sums(r2[c]);
                    const size_t e = (size_t)row * cpr + 2 * pc + c;
                    if (threadIdx.x == 0 && red) write_sums(L.red, e * RED_W, t);
                    if (DS && threadIdx.x == 0 && L.dsum) { L.dsum[2 * e] = t.dm; L.dsum[2 * e + 1] = t.dd; }
                }
            }
            return;
        }
    }
    if (rpw == 1) {
        // one (row, chunk) per workgroup: row, y and z are uniform over the workgroup
        const int cpr = cs_chunks(nx);
        const int row = b / cpr, chunk = b % cpr;
        const S *__restrict__ cur = base + (size_t)row * Q * nx;
        if (cs_aligned<S>() && nx % CS_THREADS == 0) {
            cs_aligned_chunk<S, DS, T>(L, cur, nxt, row, chunk * CS_THREADS, G, om, red, rs);
        } else {
            const int x = chunk * CS_THREADS + threadIdx.x;
            if (x < nx) cs_scatter<S, DS, T>(L, cur, nxt, row, x, G, om, red, rs);  // a partial chunk still sums below
        }
        if (red || (DS && L.dsum)) {  // uniform per workgroup
            const RowSums t = block_sums(rs);
            const size_t e = (size_t)row * cpr + chunk;
            if (threadIdx.x == 0 && red) write_sums(L.red, e * RED_W, t);
            if (DS && threadIdx.x == 0 && L.dsum) { L.dsum[2 * e] = t.dm; L.dsum[2 * e + 1] = t.dd; }
        }
    } else {
        // narrow lattice: rpw whole rows per workgroup, one partial per row summed in x order
        const int r = threadIdx.x / nx, x = threadIdx.x - r * nx;
        const int row = b * rpw + r;
        if (r < rpw && row < ny * L.nz) cs_scatter<S, DS, T>(L, base + (size_t)row * Q * nx, nxt, row, x, G, om, red, rs);
        if (red || (DS && L.dsum)) {
            __shared__ double lr[DS ? 5 : 3][CS_THREADS];
            lr[0][threadIdx.x] = rs.rho;
            lr[1][threadIdx.x] = rs.mom;
            lr[2][threadIdx.x] = rs.u;
            if constexpr (DS) {
                lr[3][threadIdx.x] = rs.dm;
                lr[4][threadIdx.x] = rs.dd;
            }
            __syncthreads();
            const int rr = b * rpw + (int)threadIdx.x;
            if ((int)threadIdx.x < rpw && rr < ny * L.nz) {
                RowSums t;
                for (int k = threadIdx.x * nx; k < (int)(threadIdx.x + 1) * nx; k++) {
                    t.rho += lr[0][k];
                    t.mom += lr[1][k];
                    t.u += lr[2][k];
                    if constexpr (DS) {
                        t.dm += lr[3][k];
                        t.dd += lr[4][k];
                    }
                }
                if (red) write_sums(L.red, (size_t)rr * RED_W, t);
                if (DS && L.dsum) { L.dsum[2 * (size_t)rr] = t.dm; L.dsum[2 * (size_t)rr + 1] = t.dd; }
            }
        }
    }
}

hipError_t launch_collide_stream(hipStream_t s, const LatArgs *lats, int nlat, int prec, const int *halt) {
    for (int i0 = 0; i0 < nlat; i0 += BATCH_MAX) {
        Batch<LatArgs> P;
        P.halt = halt;
        P.n = nlat - i0 < BATCH_MAX ? nlat - i0 : BATCH_MAX;
        bool ds = false;
        for (int i = 0; i < P.n; i++) ds = ds || lats[i0 + i].dsum != nullptr;
        int blocks = 0;
        for (int i = 0; i < BATCH_MAX; i++) {
            P.it[i] = lats[i0 + (i < P.n ? i : 0)];
            P.it[i].paired = (P.it[i].paired >> (ds ? 1 : 0)) & 1;  // the form for this launch
            P.first[i] = blocks;
            if (i < P.n) blocks += cs_blocks(P.it[i].nx, P.it[i].ny * P.it[i].nz, P.it[i].paired);
        }
        P.first[BATCH_MAX] = blocks;
        if (blocks == 0) continue;
        const dim3 g(blocks), b(CS_THREADS);
        if (prec == DNSG_PREC_FP64) {
            if (ds) hipLaunchKernelGGL((k_collide_stream<double, true>), g, b, 0, s, P);
            else hipLaunchKernelGGL((k_collide_stream<double, false>), g, b, 0, s, P);
        } else if (prec == DNSG_PREC_FP32) {
            if (ds) hipLaunchKernelGGL((k_collide_stream<float, true>), g, b, 0, s, P);
            else hipLaunchKernelGGL((k_collide_stream<float, false>), g, b, 0, s, P);
        } else {
            if (ds) hipLaunchKernelGGL((k_collide_stream<__half, true>), g, b, 0, s, P);
            else hipLaunchKernelGGL((k_collide_stream<__half, false>), g, b, 0, s, P);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---- initial state --------------------------------------------------------------
// Counter-based PRNG (splitmix64 finaliser) over (seed, grid, component, node):
// independent of the traversal order, so every node's noise is fixed by its
// global index in the reference layout [x][y][k] of that grid.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double prng_sym(uint64_t seed, int grid, int comp, uint64_t node) {
    const uint64_t b =
        mix64(mix64(seed) ^ mix64(((uint64_t)(grid & 0xff) << 56) ^ ((uint64_t)(comp & 0xff) << 48) ^ node));
    const double u = (double)(b >> 11) * 0x1.0p-53;
    return 2. * u - 1.;
}

struct InitArgs {
    void *dst;
    int nx, ny, nz;     // stored extents
    int zdf_or_zdc;     // reference z extent of this grid
    int kofs;           // reference k = stored z + kofs
    int grid;           // 0 coarse, 1 lower, 2 upper (PRNG stream)
    int kind;           // DNSG_INIT_*
    int GR, zdf, zdc, NZs;
    uint64_t seed;
    double u_max, noise;
};

template <class S>
__global__ __launch_bounds__(256) void k_init(InitArgs A) {
    const size_t n = (size_t)A.nx * A.ny * A.nz;
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int x = (int)(id % A.nx), y = (int)((id / A.nx) % A.ny), z = (int)(id / ((size_t)A.nx * A.ny));
    const int k = z + A.kofs;
    S *dst = static_cast<S *>(A.dst);
    double f[Q];
    if (A.kind == DNSG_INIT_REST) {
        unroll<0, Q>([&](auto qc) { f[decltype(qc)::value] = WQ[decltype(qc)::value]; });
    } else {
        double dist;
        if (A.grid == 0) dist = A.GR * (k - 1) + (A.zdf - 1 - 0.5);
        else if (A.grid == 1) dist = k - 0.5;
        else dist = k + A.GR * (A.zdc - 3) + A.zdf - 1. - 0.5;
        const double zeta = dist / ((double)(A.NZs - 2));
        const double amp = A.noise * A.u_max;
        const uint64_t node = ((uint64_t)x * A.ny + y) * A.zdf_or_zdc + k;
        const double u0 = 4. * A.u_max * zeta * (1. - zeta) + amp * prng_sym(A.seed, A.grid, 0, node);
        const double u1 = amp * prng_sym(A.seed, A.grid, 1, node);
        const double u2 = amp * prng_sym(A.seed, A.grid, 2, node);
        const double v = u0 * u0 + u1 * u1 + u2 * u2;
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = feq64<q>(1., u0, u1, u2, v);
        });
    }
    const size_t row = (size_t)z * A.ny + y;
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (sizeof(S) == 8) store_s((double *)dst + lat_at(q, row, x, A.nx), f[q]);
        else store_s(dst + lat_at(q, row, x, A.nx), (float)(f[q] - WQ[q]));
    });
}

hipError_t launch_init_lattice(hipStream_t s, dnsg_ctx *h, int grid, void *dst) {
    InitArgs A;
    const Grid &g = grid == 0 ? h->coarse : (grid == 1 ? h->lower : h->upper);
    A.dst = dst;
    A.nx = g.nx; A.ny = g.ny; A.nz = g.nz;
    A.zdf_or_zdc = grid == 0 ? h->d.zdc : h->d.zdf;
    A.kofs = grid == 1 ? 1 : 0;
    A.grid = grid;
    A.kind = h->p.init_kind;
    A.GR = h->d.GR; A.zdf = h->d.zdf; A.zdc = h->d.zdc; A.NZs = h->d.NZs;
    A.seed = h->p.seed; A.u_max = h->p.u_max; A.noise = h->p.noise;
    const size_t n = g.nodes();
    dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    if (h->p.precision == DNSG_PREC_FP64) hipLaunchKernelGGL(k_init<double>, grd, blk, 0, s, A);
    else if (h->p.precision == DNSG_PREC_FP32) hipLaunchKernelGGL(k_init<float>, grd, blk, 0, s, A);
    else hipLaunchKernelGGL(k_init<__half>, grd, blk, 0, s, A);
    return hipGetLastError();
}

// ---- AoS (reference layouts [x][y][k][19]) <-> SoA ----------------------------------
// gl = 0: the file layout [nx][ny][nz_aos][19]; gl = 1: the in-memory layout
// [nx+1][ny+1][nz_aos][19] whose last x-plane and y-row are ghost cells.
template <class S>
__global__ __launch_bounds__(256) void k_aos_to_soa(const double *__restrict__ aos, S *__restrict__ soa, int nx,
                                                    int ny, int nz_aos, int z0, int nz, int gl) {
    const size_t n = (size_t)nx * ny * nz * Q;
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int x = (int)(id % nx), q = (int)((id / nx) % Q);
    const size_t row = id / ((size_t)nx * Q);
    const int y = (int)(row % ny), z = (int)(row / ny);
    const double v = aos[(((size_t)x * (ny + gl) + y) * nz_aos + (z + z0)) * Q + q];
    if constexpr (sizeof(S) == 8) soa[id] = v;
    else store_s(soa + id, (float)(v - WQ[q]));
}

// every node of the host layout, ghost cells as their periodic images
template <class S>
__global__ __launch_bounds__(256) void k_soa_to_aos(const S *__restrict__ soa, double *__restrict__ aos, int nx,
                                                    int ny, int nz_aos, int z0, int nz, int gl) {
    const int mx = nx + gl, my = ny + gl;
    const size_t n = (size_t)mx * my * nz * Q;
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int x = (int)(id % mx), q = (int)((id / mx) % Q);
    const size_t r = id / ((size_t)mx * Q);
    const int y = (int)(r % my), z = (int)(r / my);
    const int xs = x < nx ? x : x - nx, ys = y < ny ? y : y - ny;
    const size_t src = lat_at(q, (size_t)z * ny + ys, xs, nx);
    double v;
    if constexpr (sizeof(S) == 8) v = soa[src];
    else v = (double)load_s(soa + src) + WQ[q];
    aos[(((size_t)x * my + y) * nz_aos + (z + z0)) * Q + q] = v;
}

hipError_t launch_aos_to_soa(hipStream_t s, const double *aos, void *soa, int nx, int ny, int nz_aos, int z0, int nz,
                             int prec, int gl) {
    const size_t n = (size_t)nx * ny * nz * Q;
    dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    if (prec == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_aos_to_soa<double>, grd, blk, 0, s, aos, (double *)soa, nx, ny, nz_aos, z0, nz, gl);
    else if (prec == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_aos_to_soa<float>, grd, blk, 0, s, aos, (float *)soa, nx, ny, nz_aos, z0, nz, gl);
    else
        hipLaunchKernelGGL(k_aos_to_soa<__half>, grd, blk, 0, s, aos, (__half *)soa, nx, ny, nz_aos, z0, nz, gl);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(hipStream_t s, const void *soa, double *aos, int nx, int ny, int nz_aos, int z0, int nz,
                             int prec, int gl) {
    const size_t n = (size_t)(nx + gl) * (ny + gl) * nz * Q;
    dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    if (prec == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_soa_to_aos<double>, grd, blk, 0, s, (const double *)soa, aos, nx, ny, nz_aos, z0, nz, gl);
    else if (prec == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_soa_to_aos<float>, grd, blk, 0, s, (const float *)soa, aos, nx, ny, nz_aos, z0, nz, gl);
    else
        hipLaunchKernelGGL(k_soa_to_aos<__half>, grd, blk, 0, s, (const __half *)soa, aos, nx, ny, nz_aos, z0, nz, gl);
    return hipGetLastError();
}

}  // namespace dnsg
