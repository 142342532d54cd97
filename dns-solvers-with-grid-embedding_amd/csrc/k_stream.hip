// k_stream.hip -- fused BGK collide + push-stream on the coarse grid and the
// wall bands, the initial-state kernel and the AoS <-> SoA transposes.
//
// One workgroup owns a chunk of CS_THREADS nodes of one (y, z) row: its reads
// are 19 contiguous vectors of the row block, its writes the same vectors of
// the 9 neighbouring rows shifted by e_x.
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

template <class S>
__global__ __launch_bounds__(CS_THREADS) void k_collide_stream(Batch<LatArgs> P) {
    const int bg = xcd_remap(blockIdx.x, gridDim.x);
    const int it = batch_find(P.first, P.n, bg);
    const LatArgs &L = P.it[it];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int cpr = (nx + CS_THREADS - 1) / CS_THREADS;  // chunks per row
    const int b = bg - P.first[it];
    const int row = b / cpr, chunk = b % cpr;
    const double *__restrict__ force = L.force;
    using T = typename ComputeOf<S>::T;
    const int y = row % ny, z = row / ny;
    const S *__restrict__ cur = static_cast<const S *>(L.cur) + (size_t)row * Q * nx;
    S *__restrict__ nxt = static_cast<S *>(L.nxt);
    const T G = (T)(L.gscale * force[0]);
    const T om = (T)L.omega;
    const bool lo_edge = (z == 0), hi_edge = (z == nz - 1);
    const int ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const bool red = L.red != nullptr;
    double r_rho = 0., r_mom = 0., r_u = 0.;
    const int x = chunk * CS_THREADS + threadIdx.x;
    if (x < nx) {  // a partial last chunk still joins the reduction below
        T f[Q];
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = load_s(cur + (size_t)q * nx + x);
        });
        if (red) {  // moments of the pre-collision state, for the fused reductions
            double d = 0., m = 0.;
            if constexpr (sizeof(S) == 8) {
                m = first_moment<0>(f);
                d = zeroth_moment(f);
            } else {
                m = (double)first_moment<0>(f);
                d = 1. + (double)zeroth_moment(f);
            }
            r_rho += d;
            r_mom += m;
            r_u += m / d;
        }
        collide_any(f, G, om);
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
                        const int8_t *wc = (EZ[q] < 0) ? L.wlo : L.whi;
                        const int sld = wc[(size_t)yt * nx + xt];
                        const int sld1 = wc[(size_t)y * nx + x];
                        if (sld == -1 && sld1 == -1)
                            store_s(nxt + lat_at(OPP[q], (size_t)row, x, nx), f[q]);
                        else
                            store_s(nxt + lat_at(MIR[q], (size_t)z * ny + yt, xt, nx), f[q]);
                    }
                    return;
                }
            }
            const int zt = z + EZ[q];
            if ((zt == 0 && L.skip_lo) || (zt == nz - 1 && L.skip_hi)) return;
            store_s(nxt + lat_at(q, (size_t)zt * ny + yt, xt, nx), f[q]);
        });
    }
    if (red) {  // uniform per workgroup
        __shared__ double lds[3][4];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        r_rho = wave_sum64(r_rho);
        r_mom = wave_sum64(r_mom);
        r_u = wave_sum64(r_u);
        if (lane == 0) { lds[0][wid] = r_rho; lds[1][wid] = r_mom; lds[2][wid] = r_u; }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int nw = (blockDim.x + 63) >> 6;
            double a = 0., bb = 0., c = 0.;
            for (int w = 0; w < nw; w++) { a += lds[0][w]; bb += lds[1][w]; c += lds[2][w]; }
            const size_t e = ((size_t)row * cpr + chunk) * RED_W;
            L.red[e + 0] = a;
            L.red[e + 1] = bb;
            L.red[e + 2] = c;
        }
    }
}

hipError_t launch_collide_stream(hipStream_t s, const LatArgs *lats, int nlat, int prec) {
    for (int i0 = 0; i0 < nlat; i0 += BATCH_MAX) {
        Batch<LatArgs> P;
        P.n = nlat - i0 < BATCH_MAX ? nlat - i0 : BATCH_MAX;
        int blocks = 0;
        for (int i = 0; i < BATCH_MAX; i++) {
            P.it[i] = lats[i0 + (i < P.n ? i : 0)];
            P.first[i] = blocks;
            if (i < P.n) blocks += P.it[i].ny * P.it[i].nz * cs_chunks(P.it[i].nx);
        }
        P.first[BATCH_MAX] = blocks;
        if (blocks == 0) continue;
        if (prec == DNSG_PREC_FP64)
            hipLaunchKernelGGL(k_collide_stream<double>, dim3(blocks), dim3(CS_THREADS), 0, s, P);
        else if (prec == DNSG_PREC_FP32)
            hipLaunchKernelGGL(k_collide_stream<float>, dim3(blocks), dim3(CS_THREADS), 0, s, P);
        else
            hipLaunchKernelGGL(k_collide_stream<__half>, dim3(blocks), dim3(CS_THREADS), 0, s, P);
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

// ---- AoS (reference file layout [x][y][k][19]) <-> SoA -----------------------------
template <class S>
__global__ __launch_bounds__(256) void k_aos_to_soa(const double *__restrict__ aos, S *__restrict__ soa, int nx,
                                                    int ny, int nz_aos, int z0, int nz) {
    const size_t n = (size_t)nx * ny * nz * Q;
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int x = (int)(id % nx), q = (int)((id / nx) % Q);
    const size_t row = id / ((size_t)nx * Q);
    const int y = (int)(row % ny), z = (int)(row / ny);
    const double v = aos[(((size_t)x * ny + y) * nz_aos + (z + z0)) * Q + q];
    if constexpr (sizeof(S) == 8) soa[id] = v;
    else store_s(soa + id, (float)(v - WQ[q]));
}

template <class S>
__global__ __launch_bounds__(256) void k_soa_to_aos(const S *__restrict__ soa, double *__restrict__ aos, int nx,
                                                    int ny, int nz_aos, int z0, int nz) {
    const size_t n = (size_t)nx * ny * nz * Q;
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int x = (int)(id % nx), q = (int)((id / nx) % Q);
    const size_t row = id / ((size_t)nx * Q);
    const int y = (int)(row % ny), z = (int)(row / ny);
    double v;
    if constexpr (sizeof(S) == 8) v = soa[id];
    else v = (double)load_s(soa + id) + WQ[q];
    aos[(((size_t)x * ny + y) * nz_aos + (z + z0)) * Q + q] = v;
}

hipError_t launch_aos_to_soa(hipStream_t s, const double *aos, void *soa, int nx, int ny, int nz_aos, int z0, int nz,
                             int prec) {
    const size_t n = (size_t)nx * ny * nz * Q;
    dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    if (prec == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_aos_to_soa<double>, grd, blk, 0, s, aos, (double *)soa, nx, ny, nz_aos, z0, nz);
    else if (prec == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_aos_to_soa<float>, grd, blk, 0, s, aos, (float *)soa, nx, ny, nz_aos, z0, nz);
    else
        hipLaunchKernelGGL(k_aos_to_soa<__half>, grd, blk, 0, s, aos, (__half *)soa, nx, ny, nz_aos, z0, nz);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(hipStream_t s, const void *soa, double *aos, int nx, int ny, int nz_aos, int z0, int nz,
                             int prec) {
    const size_t n = (size_t)nx * ny * nz * Q;
    dim3 grd((unsigned)((n + 255) / 256)), blk(256);
    if (prec == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_soa_to_aos<double>, grd, blk, 0, s, (const double *)soa, aos, nx, ny, nz_aos, z0, nz);
    else if (prec == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_soa_to_aos<float>, grd, blk, 0, s, (const float *)soa, aos, nx, ny, nz_aos, z0, nz);
    else
        hipLaunchKernelGGL(k_soa_to_aos<__half>, grd, blk, 0, s, (const __half *)soa, aos, nx, ny, nz_aos, z0, nz);
    return hipGetLastError();
}

}  // namespace dnsg
