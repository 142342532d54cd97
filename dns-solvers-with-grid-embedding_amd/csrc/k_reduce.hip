// k_reduce.hip -- FP64 reductions of the time step: bulk flux and density with
// the flux controller (encalc2.c:17-147), wall shear stress (wshearstr.c:15-168),
// the global mass check (main:1193-1250) and the plane statistics
// (statistics.c:15-300).
//
// Every reduction is two-level and deterministic: one workgroup per lattice row
// sums its row (wave64 shuffles, then LDS across the 4 waves) into a row
// partial; a single-workgroup finaliser sums the row partials in a fixed order,
// applies the reference's weights and normalisations and writes the result to
// device memory (the force the next collision reads, the diag record).  No
// atomics: the result depends only on the lattice contents, never on timing or
// on other work in flight.
#include <algorithm>

#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {

constexpr int RB = 256;  // threads per reduction workgroup
constexpr int NMOM = 5;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// block-wide sum of K values over NB threads; result valid in thread 0
template <int K, int NB = RB>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*lds)[NB / 64]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double s = wave_sum(v[k]);
        if (lane == 0) lds[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0.;
            for (int w = 0; w < NB / 64; w++) s += lds[k][w];
            v[k] = s;
        }
    }
    __syncthreads();
}

template <class S>
__device__ __forceinline__ void node_moments(const S *lat, int nx, size_t row, int x, double &dens, double &ux,
                                             double &uy, double &uz) {
    double f[Q];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (sizeof(S) == 8) f[q] = lat[lat_at(q, row, x, nx)];
        else f[q] = (double)load_s(lat + lat_at(q, row, x, nx));
    });
    ux = first_moment<0>(f);
    uy = first_moment<1>(f);
    uz = first_moment<2>(f);
    dens = zeroth_moment(f);
    if constexpr (sizeof(S) != 8) dens = 1. + dens;
}

// one lattice for the row kernels
struct RedLat {
    const void *lat;
    int nx, ny, nz;
    double *part;          // NMOM doubles per row
    const double *means;   // pass 2: [nz][4] plane means (U, V, W, rho) or null
};
struct RedSet {
    RedLat g[3];
};

// MODE 1: sums of rho, rho u_x, u_x, u_y, u_z (u = rho u / rho)
// MODE 2: sums of u'^2, v'^2, w'^2, u'w', rho'^2 about the plane means
// MODE 3: sums of u^2, v^2, w^2, u w, rho^2 (raw second moments)
template <class S, int MODE>
__global__ __launch_bounds__(RB) void k_row_moments(RedSet R) {
    const RedLat &L = R.g[blockIdx.y];
    const int row = blockIdx.x;
    if (row >= L.ny * L.nz) return;
    __shared__ double lds[NMOM][RB / 64];
    const int z = row / L.ny;
    const S *lat = static_cast<const S *>(L.lat);
    double acc[NMOM] = {0., 0., 0., 0., 0.};
    double U = 0., V = 0., Wm = 0., D = 0.;
    if (MODE == 2) {
        U = L.means[z * 4 + 0]; V = L.means[z * 4 + 1]; Wm = L.means[z * 4 + 2]; D = L.means[z * 4 + 3];
    }
    for (int x = threadIdx.x; x < L.nx; x += RB) {
        double dens, ux, uy, uz;
        node_moments(lat, L.nx, (size_t)row, x, dens, ux, uy, uz);
        if (MODE == 1) {
            acc[0] += dens; acc[1] += ux; acc[2] += ux / dens; acc[3] += uy / dens; acc[4] += uz / dens;
        } else if (MODE == 2) {
            const double uf = (ux / dens) - U, vf = (uy / dens) - V, wf = (uz / dens) - Wm, df = dens - D;
            acc[0] += uf * uf; acc[1] += vf * vf; acc[2] += wf * wf; acc[3] += uf * wf; acc[4] += df * df;
        } else {
            const double u = ux / dens, v = uy / dens, w = uz / dens;
            acc[0] += u * u; acc[1] += v * v; acc[2] += w * w; acc[3] += u * w; acc[4] += dens * dens;
        }
    }
    block_sum<NMOM>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < NMOM; k++) L.part[(size_t)row * NMOM + k] = acc[k];
}

template <int MODE>
static void row_moments(hipStream_t s, int prec, const RedSet &R, int nlat) {
    int rows = 0;
    for (int i = 0; i < nlat; i++) rows = std::max(rows, R.g[i].ny * R.g[i].nz);
    dim3 grd(rows, nlat), blk(RB);
    if (prec == DNSG_PREC_FP64) hipLaunchKernelGGL((k_row_moments<double, MODE>), grd, blk, 0, s, R);
    else if (prec == DNSG_PREC_FP32) hipLaunchKernelGGL((k_row_moments<float, MODE>), grd, blk, 0, s, R);
    else hipLaunchKernelGGL((k_row_moments<__half, MODE>), grd, blk, 0, s, R);
}

// partial-buffer layout: [coarse rows | lower rows | upper rows] x NMOM, then the
// wall-stress partials
struct PartLayout {
    size_t off_c, off_l, off_u, off_ws, total;
};
static PartLayout part_layout(const dnsg_ctx *h) {
    PartLayout P;
    const size_t rc = (size_t)h->coarse.ny * h->coarse.nz, rl = (size_t)h->lower.ny * h->lower.nz;
    P.off_c = 0;
    P.off_l = P.off_c + rc * NMOM;
    P.off_u = P.off_l + rl * NMOM;
    P.off_ws = P.off_u + rl * NMOM;
    P.total = P.off_ws + (size_t)h->lower.ny * 4;
    return P;
}

static RedSet make_set(dnsg_ctx *h, const double *means_c, const double *means_l, const double *means_u) {
    PartLayout P = part_layout(h);
    RedSet R;
    R.g[0] = {h->coarse.c(), h->coarse.nx, h->coarse.ny, h->coarse.nz, h->part + P.off_c, means_c};
    R.g[1] = {h->lower.c(), h->lower.nx, h->lower.ny, h->lower.nz, h->part + P.off_l, means_l};
    R.g[2] = {h->upper.c(), h->upper.nx, h->upper.ny, h->upper.nz, h->part + P.off_u, means_u};
    return R;
}

// ---- encalc2 ---------------------------------------------------------------------------
struct EncArgs {
    const double *pc, *pl, *pu;
    int st;            // doubles per partial: [0] sum rho, [1] sum rho u_x
    int cpr_c, cpr_f;  // partials per row (coarse, bands)
    int Yc, zdc, Yf, zds, GR, zdf, NZs, Xf, force_mode;
    double ub, g_const;
    double *force, *diag;
};

// encalc2 partials as one index space: [lower entries | coarse entries | upper entries]
struct EncSpace {
    int nl, nc, nu0, nu;  // entries in each part; nu0 = first upper entry used
    __device__ int total() const { return nl + nc + nu; }
};
__device__ __forceinline__ EncSpace enc_space(const EncArgs &A) {
    EncSpace S;
    S.nl = A.Yf * (A.zdf - A.GR - 1) * A.cpr_f;  // lower band k = 1 .. zdf-GR-1 (stored m = 0 .. zdf-GR-2)
    S.nc = A.Yc * A.zdc * A.cpr_c;                // every coarse layer
    S.nu0 = A.Yf * A.GR * A.cpr_f;                // upper band k = GR .. zdf-2 (stored m = GR .. zds-1)
    S.nu = A.Yf * A.zds * A.cpr_f - S.nu0;
    return S;
}

// v += weighted (momentum, density, coarse momentum, coarse density) of entry e
__device__ __forceinline__ void enc_entry(const EncArgs &A, const EncSpace &S, int e, double *v) {
    if (e < S.nl) {
        v[0] += A.pl[(size_t)e * A.st + 1];
        v[1] += A.pl[(size_t)e * A.st + 0];
    } else if (e < S.nl + S.nc) {
        e -= S.nl;
        const int z = (e / A.cpr_c) / A.Yc;
        const double c = (z == 0 || z == A.zdc - 1) ? 0.5 * A.GR * A.GR * A.GR : 1. * A.GR * A.GR * A.GR;
        const double m = c * A.pc[(size_t)e * A.st + 1], d = c * A.pc[(size_t)e * A.st + 0];
        v[0] += m; v[1] += d; v[2] += m; v[3] += d;
    } else {
        e = e - S.nl - S.nc + S.nu0;
        const double c = ((e / A.cpr_f) / A.Yf == A.GR) ? 0.5 : 1.;  // half weight at k = GR
        v[0] += c * A.pu[(size_t)e * A.st + 1];
        v[1] += c * A.pu[(size_t)e * A.st + 0];
    }
}

constexpr int ENC_SEG = 512;  // entries per first-stage workgroup (~240 workgroups at the Re_tau=180 size)

__global__ __launch_bounds__(RB) void k_enc_seg(EncArgs A, double *seg) {
    __shared__ double lds[4][RB / 64];
    const EncSpace S = enc_space(A);
    const int e0 = blockIdx.x * ENC_SEG, e1 = min(S.total(), e0 + ENC_SEG);
    double v[4] = {0., 0., 0., 0.};
    for (int e = e0 + threadIdx.x; e < e1; e += RB) enc_entry(A, S, e, v);
    block_sum<4>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) seg[blockIdx.x * 4 + k] = v[k];
}

// encalc2's force and log values from the sums (momentum, denst, cmomentum, cdenst)
__device__ void enc_finish(const EncArgs &A, const double *v) {
    const double vol = ((double)((A.Xf + 1 - 1.) * (A.Yf + 1 - 1.) * (A.NZs - 2.)));  // encalc2.c:109
    const double momentum = v[0] / vol, denst = v[1] / vol;
    const double G = (A.force_mode == DNSG_FORCE_FLUX) ? A.ub - momentum : A.g_const;  // :129
    A.force[0] = G;
    A.force[1] = v[2] / vol;
    A.force[2] = v[3] / vol;
    if (A.diag) {
        A.diag[1] = G;
        A.diag[2] = momentum;
        A.diag[3] = denst;
        A.diag[4] = momentum * ((A.Yf) * (A.NZs - 2)) / denst;  // :138
    }
}

__global__ __launch_bounds__(RB) void k_fin_encalc2(EncArgs A, const double *seg, int nseg) {
    __shared__ double lds[4][RB / 64];
    double v[4] = {0., 0., 0., 0.};  // momentum, denst, cmomentum, cdenst
    for (int b = threadIdx.x; b < nseg; b += RB)
        for (int k = 0; k < 4; k++) v[k] += seg[b * 4 + k];
    block_sum<4>(v, lds);
    if (threadIdx.x == 0) enc_finish(A, v);
}

static hipError_t enc_two_stage(hipStream_t s, const EncArgs &A, double *seg) {
    const int total = A.Yf * (A.zdf - A.GR - 1) * A.cpr_f + A.Yc * A.zdc * A.cpr_c + A.Yf * (A.zds - A.GR) * A.cpr_f;
    const int nseg = (total + ENC_SEG - 1) / ENC_SEG;
    hipLaunchKernelGGL(k_enc_seg, dim3(nseg), dim3(RB), 0, s, A, seg);
    hipLaunchKernelGGL(k_fin_encalc2, dim3(1), dim3(RB), 0, s, A, (const double *)seg, nseg);
    return hipGetLastError();
}

constexpr int WS_SEG = 512;  // wall-stress entries per first-stage workgroup (launch_reductions_fused)

size_t enc_seg_doubles(const dnsg_ctx *h) {
    const size_t total = (size_t)h->d.Yc * h->d.zdc * cs_chunks(h->d.Xc) +
                         2 * (size_t)h->d.Yf * (h->d.zdf - 1) * band_cpr_max(h) + (size_t)ftcd_blocks(h);
    const size_t ws = (size_t)h->d.Yf * band_cpr_max(h);
    return 4 * (total / ENC_SEG + 2 + ws / WS_SEG + 1);
}

// ---- flux controller from the writes of a step ------------------------------------------
// The state at the top of step n+1 is written during step n by: the coarse
// collide-stream (coarse layers 1..zdc-2), the injection (coarse layers 0, zdc-1),
// the last band collide-stream (every band layer except the interface layer, which
// carries no encalc2 weight).  Each writer sums its writes weighted by their
// destination layer's encalc2 weight (encalc2.c:37-105), so G = ub - <rho u> of the
// next step needs no pass over the lattices.

size_t dsum_doubles(const dnsg_ctx *h) {
    return 2 * ((size_t)h->d.Yc * h->d.zdc * cs_chunks(h->d.Xc) +
                2 * (size_t)h->d.Yf * (h->d.zdf - 1) * band_cpr_max(h));
}

void set_dest_weights(const dnsg_ctx *h, int grid, LatArgs *L) {
    const int GR = h->d.GR, zdf = h->d.zdf, zds = zdf - 1;
    if (grid == 0) { L->klo = 1; L->khi = h->d.zdc - 2; L->khalf = -1; L->kscale = 1. * GR * GR * GR; }
    else if (grid == 1) { L->klo = 0; L->khi = zdf - GR - 2; L->khalf = -1; L->kscale = 1.; }  // k = 1 .. zdf-GR-1
    else { L->klo = GR; L->khi = zds - 1; L->khalf = GR; L->kscale = 1.; }                      // k = GR .. zdf-2
}

struct FluxNextArgs {
    const double *dc, *dl, *du, *fs;  // partial pairs: coarse, lower, upper launches; injection
    int nc, nl, nu, nf;               // pairs in each
    int Xf, Yf, NZs;
    double ub;
    double *force, *enc;
    const int *halt;                  // as Batch::halt
};

__global__ __launch_bounds__(RB) void k_flux_seg(FluxNextArgs A, double *seg) {
    __shared__ double lds[4][RB / 64];
    if (halted(A.halt)) return;
    const int total = A.nc + A.nl + A.nu + A.nf;
    const int e0 = blockIdx.x * ENC_SEG, e1 = min(total, e0 + ENC_SEG);
    double v[4] = {0., 0., 0., 0.};  // momentum, density, coarse momentum, coarse density
    for (int e = e0 + threadIdx.x; e < e1; e += RB) {
        const double *p;
        bool coarse = false;
        if (e < A.nc) { p = A.dc + 2 * (size_t)e; coarse = true; }
        else if (e < A.nc + A.nl) p = A.dl + 2 * (size_t)(e - A.nc);
        else if (e < A.nc + A.nl + A.nu) p = A.du + 2 * (size_t)(e - A.nc - A.nl);
        else { p = A.fs + 2 * (size_t)(e - A.nc - A.nl - A.nu); coarse = true; }
        v[0] += p[0];
        v[1] += p[1];
        if (coarse) { v[2] += p[0]; v[3] += p[1]; }
    }
    block_sum<4>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) seg[blockIdx.x * 4 + k] = v[k];
}

__global__ __launch_bounds__(RB) void k_flux_fin(FluxNextArgs A, const double *seg, int nseg) {
    __shared__ double lds[4][RB / 64];
    if (halted(A.halt)) return;
    double v[4] = {0., 0., 0., 0.};
    for (int b = threadIdx.x; b < nseg; b += RB)
        for (int k = 0; k < 4; k++) v[k] += seg[b * 4 + k];
    block_sum<4>(v, lds);
    if (threadIdx.x == 0) {
        const double vol = ((double)((A.Xf + 1 - 1.) * (A.Yf + 1 - 1.) * (A.NZs - 2.)));  // encalc2.c:109
        const double momentum = v[0] / vol, denst = v[1] / vol;
        const double G = A.ub - momentum;                                               // :129
        A.force[0] = G;
        A.force[1] = v[2] / vol;
        A.force[2] = v[3] / vol;
        A.enc[0] = G;
        A.enc[1] = momentum;
        A.enc[2] = denst;
        A.enc[3] = momentum * ((A.Yf) * (A.NZs - 2)) / denst;  // :138
    }
}

hipError_t launch_flux_next(hipStream_t s, dnsg_ctx *h) {
    FluxNextArgs A;
    A.nc = h->d.Yc * h->d.zdc * cs_chunks(h->d.Xc);
    A.nl = A.nu = h->d.Yf * (h->d.zdf - 1) * band_cpr(h);
    A.nf = ftcd_blocks(h);
    A.dc = h->dsum;
    A.dl = A.dc + 2 * (size_t)A.nc;
    A.du = A.dl + 2 * (size_t)A.nl;
    A.fs = h->fsum;
    A.Xf = h->d.Xf; A.Yf = h->d.Yf; A.NZs = h->d.NZs; A.ub = h->p.ub;
    A.force = h->d_force; A.enc = h->enc_next;
    A.halt = h->halt();
    const int nseg = (A.nc + A.nl + A.nu + A.nf + ENC_SEG - 1) / ENC_SEG;
    hipLaunchKernelGGL(k_flux_seg, dim3(nseg), dim3(RB), 0, s, A, h->seg);
    hipLaunchKernelGGL(k_flux_fin, dim3(1), dim3(RB), 0, s, A, (const double *)h->seg, nseg);
    return hipGetLastError();
}

static EncArgs enc_args(dnsg_ctx *h, const double *pc, const double *pl, const double *pu, int st, bool fused,
                        double *diag_rec) {
    EncArgs A;
    A.pc = pc; A.pl = pl; A.pu = pu; A.st = st;
    A.cpr_c = fused ? cs_chunks(h->coarse.nx) : 1;
    A.cpr_f = fused ? band_cpr(h) : 1;
    A.Yc = h->d.Yc; A.zdc = h->d.zdc; A.Yf = h->d.Yf; A.zds = h->lower.nz; A.GR = h->d.GR; A.zdf = h->d.zdf;
    A.NZs = h->d.NZs; A.Xf = h->d.Xf; A.force_mode = h->p.force_mode; A.ub = h->p.ub; A.g_const = h->p.g_const;
    A.force = h->d_force; A.diag = diag_rec;
    return A;
}

// Bulk momentum and density of many solves at once: one workgroup per solve sums
// its whole encalc2 index space in a fixed order (deterministic) and writes
// <rho u_x>, <rho> to its item's output pair.
__global__ __launch_bounds__(RB) void k_flux_items(Batch<EncArgs> B) {
    __shared__ double lds[4][RB / 64];
    const EncArgs &A = B.it[blockIdx.x];
    const EncSpace S = enc_space(A);
    double v[4] = {0., 0., 0., 0.};
    for (int e = threadIdx.x; e < S.total(); e += RB) enc_entry(A, S, e, v);
    block_sum<4>(v, lds);
    if (threadIdx.x == 0) {
        const double vol = ((double)((A.Xf + 1 - 1.) * (A.Yf + 1 - 1.) * (A.NZs - 2.)));  // encalc2.c:109
        A.diag[0] = v[0] / vol;
        A.diag[1] = v[1] / vol;
    }
}

static EncArgs enc_args(dnsg_ctx *h, const double *pc, const double *pl, const double *pu, int st, bool fused,
                        double *diag_rec);

hipError_t launch_flux_many(hipStream_t s, dnsg_ctx *const *hs, int n, double *out, const int *slot) {
    for (int i0 = 0; i0 < n; i0 += BATCH_MAX) {
        Batch<EncArgs> B;
        B.n = std::min(BATCH_MAX, n - i0);
        for (int i = 0; i < BATCH_MAX; i++) {
            dnsg_ctx *h = hs[i0 + (i < B.n ? i : 0)];
            const double *pc = h->red;
            const double *pl = pc + (size_t)h->coarse.ny * h->coarse.nz * cs_chunks(h->coarse.nx) * RED_W;
            const double *pu = pl + (size_t)h->lower.ny * h->lower.nz * band_cpr(h) * RED_W;
            B.it[i] = enc_args(h, pc, pl, pu, RED_W, true, out + 2 * (size_t)slot[i0 + (i < B.n ? i : 0)]);
            B.first[i] = i < B.n ? i : B.n;
        }
        B.first[BATCH_MAX] = B.n;
        hipLaunchKernelGGL(k_flux_items, dim3(B.n), dim3(RB), 0, s, B);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_encalc2(hipStream_t s, dnsg_ctx *h, double *diag_rec) {
    RedSet R = make_set(h, nullptr, nullptr, nullptr);
    row_moments<1>(s, h->p.precision, R, 3);
    EncArgs A = enc_args(h, R.g[0].part, R.g[1].part, R.g[2].part, NMOM, false, diag_rec);
    return enc_two_stage(s, A, h->seg);
}


// ---- wall shear stress ---------------------------------------------------------------------
// per (x, y): lower wall  -2u(k=1) + 3u(k=2) - u(k=3), rho(k=1)
//             upper wall  -(2u(zdf-2) - 3u(zdf-3) + u(zdf-4)), rho(zdf-4)
template <class S>
__global__ __launch_bounds__(RB) void k_wshear_rows(const void *lo, const void *up, int nx, int ny, int nz,
                                                    double *part) {
    __shared__ double lds[4][RB / 64];
    const int y = blockIdx.x;
    const S *L = static_cast<const S *>(lo), *U = static_cast<const S *>(up);
    double acc[4] = {0., 0., 0., 0.};
    for (int x = threadIdx.x; x < nx; x += RB) {
        double d1, d2, d3, a1, a2, a3, t0, t1;
        node_moments(L, nx, (size_t)0 * ny + y, x, d1, a1, t0, t1);
        node_moments(L, nx, (size_t)1 * ny + y, x, d2, a2, t0, t1);
        node_moments(L, nx, (size_t)2 * ny + y, x, d3, a3, t0, t1);
        const double ux1 = a1 / d1, ux2 = a2 / d2, ux3 = a3 / d3;
        acc[0] += (-2. * ux1 + 3. * ux2 - ux3);
        acc[1] += d1;
        node_moments(U, nx, (size_t)(nz - 1) * ny + y, x, d1, a1, t0, t1);
        node_moments(U, nx, (size_t)(nz - 2) * ny + y, x, d2, a2, t0, t1);
        node_moments(U, nx, (size_t)(nz - 3) * ny + y, x, d3, a3, t0, t1);
        const double vx1 = a1 / d1, vx2 = a2 / d2, vx3 = a3 / d3;
        acc[2] += -(2. * vx1 - 3. * vx2 + vx3);
        acc[3] += d3;
    }
    block_sum<4>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[(size_t)y * 4 + k] = acc[k];
}

struct WsArgs {
    const double *part;
    int Xf, Yf, NZs, GR;
    double omega_f;
    const double *force;
    double *diag;
    double *out;        // [4]: str_up, str_down, stress, pgrad
    const double *enc;  // if set: this step's encalc2 record (G, <rho u>, <rho>, flowrate) for diag[1..4]
    int *halt;          // if set: becomes 1 when the stress is NaN (wshearstr.c:139-143 aborts there)
};

__device__ void wshear_finish(const WsArgs &A, const double *v);

__global__ __launch_bounds__(RB) void k_fin_wshear(WsArgs A) {
    __shared__ double lds[4][RB / 64];
    double v[4] = {0., 0., 0., 0.};
    for (int r = threadIdx.x; r < A.Yf; r += RB)
        for (int k = 0; k < 4; k++) v[k] += A.part[(size_t)r * 4 + k];
    block_sum<4>(v, lds);
    if (threadIdx.x == 0) wshear_finish(A, v);
}

// wall stress from the row sums of the first band substep (RED_W per row):
// sum_x (-2u1 + 3u2 - u3) = -2 S1 + 3 S2 - S3 with S_m = sum_x u_x of layer m
// One workgroup of WSB threads, four entries in flight per thread: the Re_tau=180
// bands in the two-substep layout give 8192 (row, tile column) entries per layer.
constexpr int WSB = 1024;
__global__ __launch_bounds__(WSB) void k_fin_wshear_fused(WsArgs A, const double *pl, const double *pu, int nz,
                                                          int cpr) {
    __shared__ double lds[4][WSB / 64];
    double v[4] = {0., 0., 0., 0.};
    const int Yf = A.Yf;
    auto at = [&](const double *p, int m, int y, int c, int k) { return p[(((size_t)m * Yf + y) * cpr + c) * RED_W + k]; };
#pragma unroll 4
    for (int e = threadIdx.x; e < Yf * cpr; e += WSB) {
        const int y = e / cpr, c = e % cpr;
        v[0] += (-2. * at(pl, 0, y, c, 2) + 3. * at(pl, 1, y, c, 2) - at(pl, 2, y, c, 2));
        v[1] += at(pl, 0, y, c, 0);
        v[2] += -(2. * at(pu, nz - 1, y, c, 2) - 3. * at(pu, nz - 2, y, c, 2) + at(pu, nz - 3, y, c, 2));
        v[3] += at(pu, nz - 3, y, c, 0);
    }
    block_sum<4, WSB>(v, lds);
    if (threadIdx.x == 0) {
        wshear_finish(A, v);
        if (A.enc && A.diag)
            for (int k = 0; k < 4; k++) A.diag[1 + k] = A.enc[k];
    }
}

__device__ void wshear_finish(const WsArgs &A, const double *v) {
    {
        const double mu = (1. / 6.) * (2. / A.omega_f - 1);              // wshearstr.c:35
        const double area = ((double)((A.Xf + 1 - 1.) * (A.Yf + 1 - 1.)));
        double davg = v[1] / area, stdown = v[0] / area;                   // :72-73
        stdown *= mu * davg;                                               // :81
        double davgu = v[3] / area, stup = v[2] / area;                    // :120-121
        double pgrad = A.force[0];
        pgrad = pgrad * (A.NZs - 2.) / (2. * A.GR);                        // :130-132
        stup *= mu * davgu;                                                // :134
        const double stress = 0.5 * (stup + stdown);                       // :138
        if (A.diag) { A.diag[5] = stup; A.diag[6] = stdown; A.diag[7] = stress; A.diag[8] = pgrad; }
        if (A.out) { A.out[0] = stup; A.out[1] = stdown; A.out[2] = stress; A.out[3] = pgrad; }
        if (A.halt && isnan(stress)) A.halt[0] = 1;
    }
}

hipError_t launch_wshear_fused(hipStream_t s, dnsg_ctx *h, double *diag_rec, const double *enc_rec) {
    const int cpr = band_cpr(h);
    const double *pl = h->red + (size_t)h->coarse.ny * h->coarse.nz * cs_chunks(h->coarse.nx) * RED_W;
    const double *pu = pl + (size_t)h->lower.ny * h->lower.nz * cpr * RED_W;
    WsArgs A;
    A.part = nullptr; A.Xf = h->d.Xf; A.Yf = h->d.Yf; A.NZs = h->d.NZs; A.GR = h->d.GR; A.omega_f = h->d.omega_f;
    A.force = h->d_force; A.diag = diag_rec; A.out = nullptr; A.enc = enc_rec;
    A.halt = h->halt_armed ? h->halt_flag : nullptr;
    hipLaunchKernelGGL(k_fin_wshear_fused, dim3(1), dim3(WSB), 0, s, A, pl, pu, h->lower.nz, cpr);
    return hipGetLastError();
}

// ---- encalc2 and wshearstr of one step together (constant force) ---------------------------
// Their first stages run as one launch: workgroups [0, nseg) sum encalc2's entries as
// k_enc_seg does, workgroups [nseg, nseg + nws) WS_SEG (row, tile column) entries each
// of the wall stress as k_fin_wshear_fused does.  One workgroup then finishes encalc2
// and, with the force that sets, wshearstr -- two launches where the separate
// reductions take three, one of them a single workgroup over all the wall-stress
// entries.
struct WsSegArgs {
    const double *pl, *pu;  // the bands' fused-reduction partials
    int Yf, cpr, nz;
};

__device__ __forceinline__ void ws_entry(const WsSegArgs &W, int e, double *v) {
    const int y = e / W.cpr, c = e % W.cpr;
    auto at = [&](const double *p, int m, int k) { return p[(((size_t)m * W.Yf + y) * W.cpr + c) * RED_W + k]; };
    v[0] += (-2. * at(W.pl, 0, 2) + 3. * at(W.pl, 1, 2) - at(W.pl, 2, 2));
    v[1] += at(W.pl, 0, 0);
    v[2] += -(2. * at(W.pu, W.nz - 1, 2) - 3. * at(W.pu, W.nz - 2, 2) + at(W.pu, W.nz - 3, 2));
    v[3] += at(W.pu, W.nz - 3, 0);
}

__global__ __launch_bounds__(RB) void k_red_seg(EncArgs A, WsSegArgs W, double *seg, int nseg) {
    __shared__ double lds[4][RB / 64];
    double v[4] = {0., 0., 0., 0.};
    if ((int)blockIdx.x < nseg) {  // uniform per workgroup
        const EncSpace S = enc_space(A);
        const int e0 = blockIdx.x * ENC_SEG, e1 = min(S.total(), e0 + ENC_SEG);
        for (int e = e0 + threadIdx.x; e < e1; e += RB) enc_entry(A, S, e, v);
    } else {
        const int e0 = (blockIdx.x - nseg) * WS_SEG, e1 = min(W.Yf * W.cpr, e0 + WS_SEG);
        for (int e = e0 + threadIdx.x; e < e1; e += RB) ws_entry(W, e, v);
    }
    block_sum<4>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) seg[blockIdx.x * 4 + k] = v[k];
}

__global__ __launch_bounds__(RB) void k_fin_enc_ws(EncArgs A, WsArgs Wf, const double *seg, int nseg, int nws) {
    __shared__ double lds[4][RB / 64];
    double v[4] = {0., 0., 0., 0.};
    for (int b = threadIdx.x; b < nseg; b += RB)
        for (int k = 0; k < 4; k++) v[k] += seg[b * 4 + k];
    block_sum<4>(v, lds);
    if (threadIdx.x == 0) enc_finish(A, v);  // writes the force that wshear_finish reads
    double w[4] = {0., 0., 0., 0.};
    for (int b = threadIdx.x; b < nws; b += RB)
        for (int k = 0; k < 4; k++) w[k] += seg[(nseg + b) * 4 + k];
    block_sum<4>(w, lds);
    if (threadIdx.x == 0) wshear_finish(Wf, w);
}

hipError_t launch_reductions_fused(hipStream_t s, dnsg_ctx *h, double *diag_rec) {
    const int cpr = band_cpr(h);
    const double *pc = h->red, *pl = pc + (size_t)h->coarse.ny * h->coarse.nz * cs_chunks(h->coarse.nx) * RED_W;
    const double *pu = pl + (size_t)h->lower.ny * h->lower.nz * cpr * RED_W;
    const EncArgs A = enc_args(h, pc, pl, pu, RED_W, true, diag_rec);
    const int total = A.Yf * (A.zdf - A.GR - 1) * A.cpr_f + A.Yc * A.zdc * A.cpr_c + A.Yf * (A.zds - A.GR) * A.cpr_f;
    const int nseg = (total + ENC_SEG - 1) / ENC_SEG;
    const WsSegArgs W = {pl, pu, h->d.Yf, cpr, h->lower.nz};
    const int nws = (W.Yf * W.cpr + WS_SEG - 1) / WS_SEG;
    WsArgs Wf;
    Wf.part = nullptr; Wf.Xf = h->d.Xf; Wf.Yf = h->d.Yf; Wf.NZs = h->d.NZs; Wf.GR = h->d.GR; Wf.omega_f = h->d.omega_f;
    Wf.force = h->d_force; Wf.diag = diag_rec; Wf.out = nullptr; Wf.enc = nullptr;
    Wf.halt = h->halt_armed ? h->halt_flag : nullptr;
    hipLaunchKernelGGL(k_red_seg, dim3(nseg + nws), dim3(RB), 0, s, A, W, h->seg, nseg);
    hipLaunchKernelGGL(k_fin_enc_ws, dim3(1), dim3(RB), 0, s, A, Wf, (const double *)h->seg, nseg, nws);
    return hipGetLastError();
}

hipError_t launch_wshear(hipStream_t s, dnsg_ctx *h, double *diag_rec) {
    PartLayout P = part_layout(h);
    double *part = h->part + P.off_ws;
    const Grid &l = h->lower;
    dim3 grd(l.ny), blk(RB);
    if (h->p.precision == DNSG_PREC_FP64)
        hipLaunchKernelGGL(k_wshear_rows<double>, grd, blk, 0, s, h->lower.c(), h->upper.c(), l.nx, l.ny, l.nz, part);
    else if (h->p.precision == DNSG_PREC_FP32)
        hipLaunchKernelGGL(k_wshear_rows<float>, grd, blk, 0, s, h->lower.c(), h->upper.c(), l.nx, l.ny, l.nz, part);
    else
        hipLaunchKernelGGL(k_wshear_rows<__half>, grd, blk, 0, s, h->lower.c(), h->upper.c(), l.nx, l.ny, l.nz, part);
    WsArgs A;
    A.part = part; A.Xf = h->d.Xf; A.Yf = h->d.Yf; A.NZs = h->d.NZs; A.GR = h->d.GR; A.omega_f = h->d.omega_f;
    A.force = h->d_force; A.diag = diag_rec; A.out = nullptr; A.enc = nullptr; A.halt = nullptr;
    hipLaunchKernelGGL(k_fin_wshear, dim3(1), dim3(RB), 0, s, A);
    return hipGetLastError();
}

// ---- mass check ---------------------------------------------------------------------------
struct MassArgs {
    const double *pc, *pl, *pu;
    int Yc, zdc, Yf, zds, GR, Xf, NZs;
    double wall_mass;
    double *out;
};

__global__ __launch_bounds__(RB) void k_fin_mass(MassArgs A) {
    __shared__ double lds[1][RB / 64];
    double v[1] = {0.};
    const double g3 = 1. * A.GR * A.GR * A.GR, g3h = 0.5 * A.GR * A.GR * A.GR;
    const int nb = A.Yf * A.zds;
    // lower band k = 1 .. zdf-1 (stored m = 0 .. zds-1), half at the interface k = zdf-1
    for (int r = threadIdx.x; r < nb; r += RB) v[0] += ((r / A.Yf == A.zds - 1) ? 0.5 : 1.) * A.pl[r * NMOM];
    // coarse k = 1 .. zdc-2, GR^3 (half at k = 1, zdc-2)
    for (int r = A.Yc + threadIdx.x; r < A.Yc * (A.zdc - 1); r += RB) {
        const int z = r / A.Yc;
        v[0] += ((z == 1 || z == A.zdc - 2) ? g3h : g3) * A.pc[r * NMOM];
    }
    // upper band k = 0 .. zdf-2, half at the interface k = 0
    for (int r = threadIdx.x; r < nb; r += RB) v[0] += ((r / A.Yf == 0) ? 0.5 : 1.) * A.pu[r * NMOM];
    block_sum<1>(v, lds);
    if (threadIdx.x == 0)
        A.out[0] = (v[0] + A.wall_mass) / ((double)((A.Xf + 1 - 1.) * (A.Yf + 1 - 1.) * (A.NZs - 2.)));
}

hipError_t launch_mass(hipStream_t s, dnsg_ctx *h, double *out) {
    RedSet R = make_set(h, nullptr, nullptr, nullptr);
    row_moments<1>(s, h->p.precision, R, 3);
    MassArgs A;
    A.pc = R.g[0].part; A.pl = R.g[1].part; A.pu = R.g[2].part;
    A.Yc = h->d.Yc; A.zdc = h->d.zdc; A.Yf = h->d.Yf; A.zds = h->lower.nz; A.GR = h->d.GR; A.Xf = h->d.Xf;
    A.NZs = h->d.NZs; A.wall_mass = h->wall_mass; A.out = out;
    hipLaunchKernelGGL(k_fin_mass, dim3(1), dim3(RB), 0, s, A);
    return hipGetLastError();
}

// ---- plane statistics ------------------------------------------------------------------------
// profile slots: lower k = 1..zdf-1 -> k; coarse k = 2..zdc-3 -> zdf-2+k;
// upper k = 0..zdf-2 -> zdf+zdc-4+k (statistics.c:63, :97, :131)
struct PlaneArgs {
    const double *pc, *pl, *pu;
    int Yc, Yf, zdc, zdf, NZ;
    double nc, nf;      // nodes per plane (Xc*Yc, Xf*Yf)
    int pass;           // 1: means, 2: fluctuations, 3: raw second moments
    double *stats;      // [9][NZ] in the turb-fld order
    double *means_c, *means_l, *means_u;  // [nz][4]
};

__global__ __launch_bounds__(RB) void k_fin_planes(PlaneArgs A) {
    __shared__ double lds[NMOM][RB / 64];
    const int slot = blockIdx.x;
    const double *part = nullptr;
    int ny = 0, z = -1;
    double nn = 1.;
    double *mean_out = nullptr;
    if (slot >= 1 && slot <= A.zdf - 1) { part = A.pl; ny = A.Yf; z = slot - 1; nn = A.nf; mean_out = A.means_l; }
    else if (slot >= A.zdf && slot <= A.zdf + A.zdc - 5) {
        part = A.pc; ny = A.Yc; z = slot - (A.zdf - 2); nn = A.nc; mean_out = A.means_c;
    } else if (slot >= A.zdf + A.zdc - 4 && slot <= 2 * A.zdf + A.zdc - 6) {
        part = A.pu; ny = A.Yf; z = slot - (A.zdf + A.zdc - 4); nn = A.nf; mean_out = A.means_u;
    }
    double v[NMOM] = {0., 0., 0., 0., 0.};
    if (part)
        for (int y = threadIdx.x; y < ny; y += RB)
            for (int k = 0; k < NMOM; k++) v[k] += part[((size_t)z * ny + y) * NMOM + k];
    block_sum<NMOM>(v, lds);
    if (threadIdx.x != 0) return;
    double *st = A.stats;
    const int NZ = A.NZ;
    if (A.pass == 1) {
        // v = (rho, rho u_x, u, v, w): means U, V, W, rho
        const double U = part ? v[2] / nn : 0., V = part ? v[3] / nn : 0., W = part ? v[4] / nn : 0.,
                     D = part ? v[0] / nn : 0.;
        st[0 * NZ + slot] = U; st[1 * NZ + slot] = V; st[2 * NZ + slot] = W; st[7 * NZ + slot] = D;
        if (mean_out) { mean_out[z * 4 + 0] = U; mean_out[z * 4 + 1] = V; mean_out[z * 4 + 2] = W; mean_out[z * 4 + 3] = D; }
    } else if (A.pass == 2) {
        st[3 * NZ + slot] = sqrt(part ? v[0] / nn : 0.);
        st[4 * NZ + slot] = sqrt(part ? v[1] / nn : 0.);
        st[5 * NZ + slot] = sqrt(part ? v[2] / nn : 0.);
        st[6 * NZ + slot] = part ? v[3] / nn : 0.;
        st[8 * NZ + slot] = sqrt(part ? v[4] / nn : 0.);
    } else {
        // raw second moments <u^2>, <v^2>, <w^2>, <uw>, <rho^2> into rows 3..6, 8
        st[3 * NZ + slot] = part ? v[0] / nn : 0.;
        st[4 * NZ + slot] = part ? v[1] / nn : 0.;
        st[5 * NZ + slot] = part ? v[2] / nn : 0.;
        st[6 * NZ + slot] = part ? v[3] / nn : 0.;
        st[8 * NZ + slot] = part ? v[4] / nn : 0.;
    }
}

static PlaneArgs plane_args(dnsg_ctx *h, const RedSet &R, int pass, double *stats, double *mc, double *ml,
                            double *mu) {
    PlaneArgs A;
    A.pc = R.g[0].part; A.pl = R.g[1].part; A.pu = R.g[2].part;
    A.Yc = h->d.Yc; A.Yf = h->d.Yf; A.zdc = h->d.zdc; A.zdf = h->d.zdf; A.NZ = h->d.NZ;
    A.nc = (double)h->d.Xc * h->d.Yc; A.nf = (double)h->d.Xf * h->d.Yf;
    A.pass = pass; A.stats = stats; A.means_c = mc; A.means_l = ml; A.means_u = mu;
    return A;
}

// out9nz receives the 9 x NZ profiles (device pointer)
hipError_t launch_statistics(hipStream_t s, dnsg_ctx *h, double *out9nz) {
    double *mc = h->d_red, *ml = mc + 4 * h->coarse.nz, *mu = ml + 4 * h->lower.nz;
    RedSet R1 = make_set(h, nullptr, nullptr, nullptr);
    row_moments<1>(s, h->p.precision, R1, 3);
    PlaneArgs A = plane_args(h, R1, 1, out9nz, mc, ml, mu);
    hipLaunchKernelGGL(k_fin_planes, dim3(h->d.NZ), dim3(RB), 0, s, A);
    RedSet R2 = make_set(h, mc, ml, mu);
    row_moments<2>(s, h->p.precision, R2, 3);
    A.pass = 2;
    hipLaunchKernelGGL(k_fin_planes, dim3(h->d.NZ), dim3(RB), 0, s, A);
    return hipGetLastError();
}

// Running averages: per snapshot, fold the exact instantaneous profiles
// (two-pass, central moments) into 14 x NZ accumulators
//   0..2  U, V, W        3..5  u_rms^2, v_rms^2, w_rms^2   6  u'w'
//   7     rho - 1        8     rho_rms^2
//   9..11 U^2, V^2, W^2  12    U W                       13  (rho - 1)^2
// so that the time average of a second moment is the mean plane variance
// plus the variance of the plane mean (law of total variance), with the
// density kept as a deviation from 1 to avoid cancellation.
__global__ void k_fold_stats(double *acc, const double *st, int NZ) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NZ) return;
    const double U = st[0 * NZ + k], V = st[1 * NZ + k], W = st[2 * NZ + k], D = st[7 * NZ + k] - 1.;
    const double ur = st[3 * NZ + k], vr = st[4 * NZ + k], wr = st[5 * NZ + k], dr = st[8 * NZ + k];
    acc[0 * NZ + k] += U;  acc[1 * NZ + k] += V;  acc[2 * NZ + k] += W;
    acc[3 * NZ + k] += ur * ur; acc[4 * NZ + k] += vr * vr; acc[5 * NZ + k] += wr * wr;
    acc[6 * NZ + k] += st[6 * NZ + k];
    acc[7 * NZ + k] += D; acc[8 * NZ + k] += dr * dr;
    acc[9 * NZ + k] += U * U; acc[10 * NZ + k] += V * V; acc[11 * NZ + k] += W * W;
    acc[12 * NZ + k] += U * W; acc[13 * NZ + k] += D * D;
}

hipError_t launch_fold_stats(hipStream_t s, double *acc, const double *st, int NZ) {
    hipLaunchKernelGGL(k_fold_stats, dim3((NZ + 255) / 256), dim3(256), 0, s, acc, st, NZ);
    return hipGetLastError();
}

size_t partial_doubles(const dnsg_ctx *h) { return part_layout(h).total; }

}  // namespace dnsg
