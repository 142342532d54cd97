// k_band2.hip -- both band substeps of a gratio-2 coarse step in one pass over the
// band lattices.
//
// With gratio 2 a coarse step runs xfcomputations twice on each band, with the
// coarse->fine fill of the interface layer between them (main:1347-1365), so a
// two-launch step reads and writes every band value twice.  The first substep's
// result feeds only the second: the first fill writes just the interface layer,
// which the first substep does not store (it is rewritten by the transfer), and the
// fill reads only the interface planes.  So the fill runs first, into the output
// lattice's interface layer, and one workgroup then carries a tile of the band
// through both substeps, keeping the intermediate state on chip:
//   * a tile is 64 x-consecutive columns (one wave per row) by 4 rows (FP64) or 8
//     rows (FP32, FP16), all layers;
//   * the workgroup marches through the layers from the interface towards the wall.
//     At layer t it loads and collides the first substep's input on the tile and a
//     one-node ring around it (the rows above and below, one wave each, and the
//     columns left and right, one extra wave); then it gathers the first substep's
//     streamed values of layer t-1 -- along a row by wave shuffles, across rows and
//     from the ring columns through LDS, across layers from what earlier
//     iterations kept -- collides them again (the second substep) and pushes the
//     result to its targets in the output lattice with the wall rules;
//   * the interface layer (t = 0) takes the first fill's values from the output
//     lattice; the layer next to the wall (t = nz-1) takes its values coming from
//     the wall by the wall rules in pull form: for a direction b leaving the wall,
//     halfway bounce-back of the node's own opp(b) value when the wall cells at
//     (x, y) and (x, y) - e_b (in plane) are both no-slip (-1), else specular
//     reflection of mirror(b) from the node (x, y) - e_b, exactly the writes of
//     xfcomputations.c:230-239 seen from the receiving node.
// Every collision and store is the two-launch path's on the same inputs, so the
// result is bit for bit the same.  HBM traffic per band node and coarse step drops
// from 2 x (19 reads + 19 writes) to 19 reads (+ the ring, largely from L2) and 19
// writes.
//
// The same launch sums the first substep's pre-collision moments per (row, tile
// column) for encalc2 / wshearstr (LatArgs::red) and, in flux mode, the
// destination-weighted sums of the second substep's writes (LatArgs::dsum), both in
// the 64-column partial layout (B2_TX).
#include <algorithm>

#include "internal.hpp"
#include "lattice.hpp"
#include "stream_common.hpp"

namespace dnsg {

namespace {

constexpr int TX = B2_TX;          // columns per tile = lanes per row wave
constexpr int CW = TX + 2;         // LDS columns: left ring column, the tile's 64, right ring column
// Output rows per tile.  The workgroup has a row wave per tile row, one for the ring
// row above and one for the ring row below, and one wave for the ring columns; a CU
// holds one workgroup (LDS).  FP64 needs 240 VGPRs, two waves per SIMD: 4 rows, 7
// waves.  FP32 and FP16 need 160, three waves per SIMD: 8 rows, 11 waves (9 rows,
// 12 waves, measured slower).
template <class S>
constexpr int tile_rows() { return sizeof(S) == 8 ? 4 : 8; }
template <class S>
constexpr int row_waves() { return tile_rows<S>() + 2; }
template <class S>
constexpr int wg_threads() { return 64 * (row_waves<S>() + 1); }

// March coordinate t: 0 at the interface layer, nz-1 next to the wall.  A direction
// moves t by ez for the upper band (interface = stored layer 0) and by -ez for the
// lower band (interface = stored layer nz-1).
template <int SIDE>
constexpr int ezt(int q) { return SIDE ? EZ[q] : -EZ[q]; }
// class of a direction: 0 along the layer, 1 toward the wall, 2 toward the interface
template <int SIDE>
constexpr int cls(int q) { return ezt<SIDE>(q) == 0 ? 0 : (ezt<SIDE>(q) > 0 ? 1 : 2); }
// directions with e_y = 0, register slot within their class: 0 e_x = 0, 1 e_x = +1, 2 e_x = -1
constexpr int xi(int q) { return EX[q] == 0 ? 0 : (EX[q] > 0 ? 1 : 2); }
// directions with e_y != 0, LDS slot within their class: along the layer q = 3, 4, 7,
// 8, 9, 10 -> 0..5; out of the layer e_y = +1 -> 0, e_y = -1 -> 1
constexpr int yi(int q) {
    return q == 3 ? 0 : q == 4 ? 1 : q == 7 ? 2 : q == 8 ? 3 : q == 9 ? 4 : q == 10 ? 5 : (EY[q] > 0 ? 0 : 1);
}

// Post-collision values of the first substep that cross waves.  y*: directions with
// e_y != 0, by row and column; h*: ring columns, e_y = 0 directions pointing into
// the tile (left column: e_x = +1, right column: e_x = -1), by row.  Layer slots:
// along the layer t & 1 (read one iteration later), toward the wall t % 3 (read two
// iterations later), toward the interface one (read in the same iteration); two
// barriers per iteration separate the writes of a layer from the reads of older ones.
template <class T, int NR>
struct Lds {
    T yf[2][6][NR][CW];
    T yu[3][2][NR][CW];
    T yd[2][NR][CW];
    T hf[2][NR][2];
    T hu[3][NR][2];
    T hd[NR][2];
    int8_t wcl[NR][CW];  // the tile's wall codes, ring included
};

// wall codes of scatter_push from the tile's LDS copy: the node at row wave r, LDS
// column col, and its in-plane neighbours
struct WallCodesTile {
    const int8_t (*w)[CW];
    int r, col;
    __device__ __forceinline__ int operator()(int, int dx, int dy, int, int) const { return w[r + dy][col + dx]; }
};

template <class S>
__device__ __forceinline__ typename ComputeOf<S>::T rest_value(int q) {
    // a harmless stand-in for lanes outside the lattice (rho = 1, u = 0)
    if constexpr (sizeof(S) == 8) return WQ[q];
    else return 0.f;
}

template <class S, int SIDE, bool DS>
__device__ __forceinline__ void band2_tile(const LatArgs &L, int tile,
                                           Lds<typename ComputeOf<S>::T, row_waves<S>()> &sm) {
    using T = typename ComputeOf<S>::T;
    constexpr int TY = tile_rows<S>(), NR = row_waves<S>();
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int tilesx = (nx + TX - 1) / TX;
    const int tx = tile % tilesx, ty = tile / tilesx;
    const int x0 = tx * TX, y0 = ty * TY;
    const int w = min(TX, nx - x0), th = min(TY, ny - y0), nrows = th + 2;
    // the wave index as a scalar, so that the compiler sees branches on it as uniform
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int cpr = tilesx;  // partials per (layer, row)
    const S *__restrict__ cur = static_cast<const S *>(L.cur);
    S *__restrict__ nxt = static_cast<S *>(L.nxt);
    const T G = (T)(L.gscale * L.force[0]);
    const T om = (T)L.omega;
    const int8_t *__restrict__ wc = SIDE ? L.whi : L.wlo;
    // this thread's node: row wave r, column x (ring wave: ring row r, left or right column)
    const bool ringw = wave == NR;
    const int r = ringw ? (lane >> 1) : wave;
    const int sd = lane & 1;
    const bool rowok = r < nrows && (!ringw || lane < 2 * NR);
    const int y = ((y0 - 1 + r) % ny + ny) % ny;
    const int x = ringw ? (sd ? (x0 + w) % nx : (x0 - 1 + nx) % nx) : x0 + lane;
    const bool act = rowok && (ringw || lane < w);
    const int col = ringw ? (sd ? w + 1 : 0) : lane + 1;
    const bool core = !ringw && r >= 1 && r <= th;
    // in-row values kept across iterations: along the layer (layer t-1, layer t),
    // toward the wall (t-2, t-1, t), toward the interface (t)
    T fprv[3] = {0, 0, 0}, fcur[3] = {0, 0, 0};
    T uprv2[3] = {0, 0, 0}, uprv[3] = {0, 0, 0}, ucur[3] = {0, 0, 0};
    T dcur[3] = {0, 0, 0};

    // gather and collide the second substep at layer t into g (core row waves only)
    auto second = [&](int t, bool wall, T *g) __attribute__((always_inline)) {
        const int m = SIDE ? t : nz - 1 - t;
        const size_t row = (size_t)m * ny + y;
        if (t == 0) {  // the interface layer: the first fill's values
            // the addresses are formed here, not kept from the kernel's start: they are
            // loop-invariant, and holding 19 of them through the march spills registers
            const S *fill = nxt + lat_at(0, row, x, nx);
            asm volatile("" : "+v"(fill));
            unroll<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                g[q] = act ? (T)load_s(fill + (size_t)q * nx) : rest_value<S>(q);
            });
        } else {
            const int sf = t & 1, su = (t - 1) % 3, sw = t % 3;
            unroll<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int c = cls<SIDE>(q);
                if constexpr (c == 2) {
                    if (wall) {  // from the wall: the wall rules (below)
                        g[q] = T(0);
                        return;
                    }
                }
                if constexpr (EY[q] == 0) {
                    constexpr int k = xi(q);
                    const T v = c == 0 ? fprv[k] : (c == 1 ? uprv2[k] : dcur[k]);
                    if constexpr (EX[q] == 0) {
                        g[q] = v;
                    } else if constexpr (EX[q] > 0) {
                        const T s = __shfl_up(v, 1, 64);
                        const T hv = c == 0 ? sm.hf[sf][r][0] : (c == 1 ? sm.hu[su][r][0] : sm.hd[r][0]);
                        g[q] = lane == 0 ? hv : s;
                    } else {
                        const T s = __shfl_down(v, 1, 64);
                        const T hv = c == 0 ? sm.hf[sf][r][1] : (c == 1 ? sm.hu[su][r][1] : sm.hd[r][1]);
                        g[q] = lane == w - 1 ? hv : s;
                    }
                } else {
                    constexpr int j = yi(q);
                    const int rs = r - EY[q], cs = lane + 1 - EX[q];
                    if constexpr (c == 0) g[q] = sm.yf[sf][j][rs][cs];
                    else if constexpr (c == 1) g[q] = sm.yu[su][j][rs][cs];
                    else g[q] = sm.yd[j][rs][cs];
                }
            });
            if (wall) {
                // directions leaving the wall (toward the interface) at the layer next to
                // it, from that layer's toward-the-wall values (layer t, slot sw, uprv)
                const int own = act ? (int)sm.wcl[r][col] : 0;
                unroll<0, Q>([&](auto qc) {
                    constexpr int b = decltype(qc)::value;
                    if constexpr (cls<SIDE>(b) == 2) {
                        constexpr int ebx = EX[b], eby = EY[b];
                        const bool bb = act && own == -1 && sm.wcl[r - eby][col - ebx] == -1;
                        T vbb, vsp;
                        if constexpr (eby == 0) {
                            vbb = uprv[xi(OPP[b])];
                            const T v = uprv[xi(MIR[b])];
                            if constexpr (ebx == 0) {
                                vsp = v;
                            } else if constexpr (ebx > 0) {
                                const T s = __shfl_up(v, 1, 64);
                                vsp = lane == 0 ? sm.hu[sw][r][0] : s;
                            } else {
                                const T s = __shfl_down(v, 1, 64);
                                vsp = lane == w - 1 ? sm.hu[sw][r][1] : s;
                            }
                        } else {
                            vbb = sm.yu[sw][yi(OPP[b])][r][lane + 1];
                            vsp = sm.yu[sw][yi(MIR[b])][r - eby][lane + 1 - ebx];
                        }
                        g[b] = bb ? vbb : vsp;
                    }
                });
            }
        }
        collide_any(g, G, om);
    };
    // push the second substep's result at layer t to its targets
    auto push = [&](int t, const T *g) __attribute__((always_inline)) {
        const int m = SIDE ? t : nz - 1 - t;
        const size_t row = (size_t)m * ny + y;
        RowSums rs;
        if (act) scatter_push<S, DS>(L, nxt, (int)row, x, g, rs, WallCodesTile{sm.wcl, r, col});
        if (DS && L.dsum) {
            const double dm = wave_sum64(rs.dm), dd = wave_sum64(rs.dd);
            if (lane == 0) {
                const size_t e = row * cpr + tx;
                L.dsum[2 * e] = dm;
                L.dsum[2 * e + 1] = dd;
            }
        }
    };

    // the first substep's input of a layer, loaded one layer ahead: the loads of layer
    // t+1 are issued at the top of iteration t and waited for just before the second
    // substep's stores (below), so they overlap both collisions.  The loads are
    // unconditional, so no join between lanes or waves waits for them: lanes without
    // a node read the tile's first ring node, whose cache lines lane 0 of the first
    // row wave fetches anyway, and take rest values at use.
    const int yl = act ? y : ((y0 - 1) % ny + ny) % ny, xl = act ? x : x0;
    T fin[Q];
    auto load_layer = [&](int t) __attribute__((always_inline)) {
        const size_t row = (size_t)(SIDE ? t : nz - 1 - t) * ny + yl;
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            fin[q] = (T)load_s(cur + lat_at(q, row, xl, nx));
        });
    };
    // the tile's wall codes, read at the layer next to the wall
    if (act) sm.wcl[r][col] = wc[(size_t)y * nx + x];
    load_layer(0);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // the first layer is waited for on entry
    for (int t = 0; t < nz; t++) {
        // ---- first substep at layer t: load, sum, collide, share
        const int m = SIDE ? t : nz - 1 - t;
        const size_t row = (size_t)m * ny + y;
        T f[Q];
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = act ? fin[q] : rest_value<S>(q);
        });
        if (t + 1 < nz) load_layer(t + 1);
        if (rowok) {
            if (core && L.red) {
                RowSums rs;
                if (act) rs.add<S>(f);
                const double a = wave_sum64(rs.rho), b = wave_sum64(rs.mom), c = wave_sum64(rs.u);
                if (lane == 0) {
                    double *p = L.red + (row * cpr + tx) * RED_W;
                    p[0] = a;
                    p[1] = b;
                    p[2] = c;
                }
            }
            collide_any(f, G, om);
            if constexpr (sizeof(S) == 2) {
                // FP16 storage: one launch per substep stores this result in FP16 and the
                // second substep reads it back; round it the same way here
                unroll<0, Q>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    f[q] = __half2float(__float2half(f[q]));
                });
            }
            const int sf = t & 1, su = t % 3;
            unroll<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int c = cls<SIDE>(q);
                if constexpr (EY[q] != 0) {
                    // only nodes of the lattice write: a row wave's lanes past a partial
                    // tile's last column would overwrite the right ring column
                    constexpr int j = yi(q);
                    if (act) {
                        if constexpr (c == 0) sm.yf[sf][j][r][col] = f[q];
                        else if constexpr (c == 1) sm.yu[su][j][r][col] = f[q];
                        else sm.yd[j][r][col] = f[q];
                    }
                } else {
                    constexpr int k = xi(q);
                    if (ringw) {
                        // a ring column passes the direction pointing into the tile
                        if constexpr (k != 0) {
                            if (act && sd == (k == 1 ? 0 : 1)) {
                                if constexpr (c == 0) sm.hf[sf][r][sd] = f[q];
                                else if constexpr (c == 1) sm.hu[su][r][sd] = f[q];
                                else sm.hd[r][sd] = f[q];
                            }
                        }
                    } else {
                        if constexpr (c == 0) fcur[k] = f[q];
                        else if constexpr (c == 1) ucur[k] = f[q];
                        else dcur[k] = f[q];
                    }
                }
            });
        }
        __syncthreads();
        // ---- second substep at layer t-1: gather and collide, then (after the barrier
        // that frees the layer slots for the next iteration) push
        T g[Q];
        if (t >= 1 && core) second(t - 1, false, g);
        // wait here, on every wave, for every vector memory operation in flight -- the
        // next layer's loads and the previous layer's stores -- so that the pushes below
        // are the only ones outstanding when the next iteration first uses the loaded
        // values (the compiler's own wait there would cover them too).  The wait sits
        // beside the barrier, in code that every path runs.
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), no wait on expcnt / lgkmcnt
        __syncthreads();
        if (t >= 1 && core) push(t - 1, g);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            fprv[k] = fcur[k];
            uprv2[k] = uprv[k];
            uprv[k] = ucur[k];
        }
    }
    // ---- second substep at the layer next to the wall
    if (core) {
        T g[Q];
        second(nz - 1, true, g);
        push(nz - 1, g);
    }
}

template <class S, bool DS>
__global__ __launch_bounds__(wg_threads<S>()) void k_band2(Batch<LatArgs> P) {
    using T = typename ComputeOf<S>::T;
    __shared__ Lds<T, row_waves<S>()> sm;
    if (halted(P.halt)) return;
    const int bg = xcd_remap(blockIdx.x, gridDim.x);
    const int it = batch_find(P.first, P.n, bg);
    const LatArgs &L = P.it[it];
    const int tile = bg - P.first[it];
    if (L.lo_kind == BOUND_WALL) band2_tile<S, 0, DS>(L, tile, sm);
    else band2_tile<S, 1, DS>(L, tile, sm);
}

}  // namespace

int band2_tiles(int nx, int ny, int prec) {
    const int ty = prec == DNSG_PREC_FP64 ? tile_rows<double>() : tile_rows<float>();
    return ((nx + TX - 1) / TX) * ((ny + ty - 1) / ty);
}

hipError_t launch_band2(hipStream_t s, const LatArgs *bands, int prec, const int *halt) {
    Batch<LatArgs> P;
    P.halt = halt;
    P.n = 2;
    int blocks = 0;
    for (int i = 0; i < BATCH_MAX; i++) {
        P.it[i] = bands[i < 2 ? i : 0];
        P.first[i] = blocks;
        if (i < 2) blocks += band2_tiles(bands[i].nx, bands[i].ny, prec);
    }
    P.first[BATCH_MAX] = blocks;
    const bool ds = bands[0].dsum != nullptr || bands[1].dsum != nullptr;
    const dim3 g(blocks), b(prec == DNSG_PREC_FP64 ? wg_threads<double>() : wg_threads<float>());
    if (prec == DNSG_PREC_FP64) {
        if (ds) hipLaunchKernelGGL((k_band2<double, true>), g, b, 0, s, P);
        else hipLaunchKernelGGL((k_band2<double, false>), g, b, 0, s, P);
    } else if (prec == DNSG_PREC_FP32) {
        if (ds) hipLaunchKernelGGL((k_band2<float, true>), g, b, 0, s, P);
        else hipLaunchKernelGGL((k_band2<float, false>), g, b, 0, s, P);
    } else {
        if (ds) hipLaunchKernelGGL((k_band2<__half, true>), g, b, 0, s, P);
        else hipLaunchKernelGGL((k_band2<__half, false>), g, b, 0, s, P);
    }
    return hipGetLastError();
}

}  // namespace dnsg
