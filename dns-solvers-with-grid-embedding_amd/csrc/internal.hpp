// internal.hpp -- device-side state of one solver handle and the kernel launchers.
//
// Layout in HBM (row-interleaved structure of arrays, x fastest): the 19
// direction vectors of one (z, y) row are stored back to back,
//   lattice g, direction q, stored layer z, row y, column x ->
//       base[g] + ((z*ny + y)*19 + q)*nx + x
// so one workgroup's reads are a single contiguous 19*nx block and its writes
// fall in the 9 neighbouring rows' blocks.
// The coarse grid stores all zdc layers.  A band stores its zds fluid layers
// only (the wall layer is never collided or streamed and is kept on the host):
//   lower band: stored layer m <-> reference k = m+1  (wall below m = 0,
//               interface layer m = zds-1)
//   upper band: stored layer m <-> reference k = m    (interface layer m = 0,
//               wall above m = zds-1)
// Interface planes are [23][Yc][Xc] (rho, u, f_neq[19]).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dnsg.h"

namespace dnsg {

enum BoundKind { BOUND_DROP = 0, BOUND_WALL = 1 };

// one lattice as the streaming kernels see it
// element (q, row = z*ny + y, x) of a lattice with row length nx
__host__ __device__ __forceinline__ size_t lat_at(int q, size_t row, int x, int nx) {
    return (row * 19 + (size_t)q) * (size_t)nx + (size_t)x;
}

struct LatArgs {
    const void *cur;   // pre-collision state (read)
    void *nxt;         // post-stream state (written)
    int nx, ny, nz;
    int lo_kind, hi_kind;  // below z = 0 / above z = nz-1
    const int8_t *wlo;     // wall codes [ny][nx] (-1 solid, -2 gas) of the wall below
    const int8_t *whi;     // ... and above
    double gscale;         // G = gscale * force (1 for coarse, 1/GR for the bands)
    double omega;          // 1/tau
    int skip_lo, skip_hi;  // do not write into layer 0 / nz-1 (rewritten by the interface)
    double *red;           // if set: per-row sums of the pre-collision state, RED_W per row
    const double *force;   // V.force[0] of the solve this lattice belongs to
    // if set: per-(row, chunk) sums over the values this launch writes, each weighted
    // by the encalc2 weight of its destination layer: [sum w e_x f, sum w f].  Layers
    // klo..khi carry kscale (khalf: half of it), all others 0.
    double *dsum;
    int klo, khi, khalf;
    int paired;  // the paired collide-stream form (cs_paired): bit 0 in a plain launch, bit 1 in a launch with dsum
    double kscale;
};

// Several independent work items in one launch: item i owns the workgroups
// [first[i], first[i+1]) of a one-dimensional grid.  The table travels in the
// kernel arguments, so a launch needs no device-side setup.
constexpr int BATCH_MAX = 30;
template <class A>
struct Batch {
    static_assert(sizeof(A) * BATCH_MAX + 4 * (BATCH_MAX + 2) + 8 <= 4096, "batch table exceeds the kernel-argument limit");
    A it[BATCH_MAX];
    int first[BATCH_MAX + 1];
    int n;
    // if set: the launch does nothing once *halt != 0 (a NaN wall stress earlier in
    // the same queued sequence of steps)
    const int *halt = nullptr;
};
// true when the launch belongs to a halted step sequence (uniform per launch)
__device__ __forceinline__ bool halted(const int *halt) { return halt != nullptr && *halt != 0; }
__device__ __forceinline__ int batch_find(const int *first, int n, int b) {
    int i = 0;
    while (i + 1 < n && b >= first[i + 1]) i++;
    return i;
}

// per-(row, chunk) sums written by the collide-stream kernels when LatArgs::red
// is set: sum rho, sum rho u_x, sum u_x (u_x = rho u_x / rho)
constexpr int RED_W = 3;
// threads (= nodes) per collide-stream workgroup and chunks per lattice row
#ifndef DNSG_CS_THREADS
#define DNSG_CS_THREADS 256
#endif
constexpr int CS_THREADS = DNSG_CS_THREADS;
__host__ __device__ __forceinline__ int cs_chunks(int nx) { return (nx + CS_THREADS - 1) / CS_THREADS; }
// lattices narrower than a workgroup: each workgroup takes this many whole rows
__host__ __device__ __forceinline__ int cs_rows_per_wg(int nx) { return nx >= CS_THREADS ? 1 : CS_THREADS / nx; }
// Paired collide-stream form (mask bit 0 FP64, bit 1 FP32, bit 2 FP16): rows that are
// a multiple of 2 CS_THREADS nodes are taken two chunks per workgroup, two adjacent
// nodes per thread, with 2-value loads and stores.  Measured on the Re_tau=180
// lattices: FP32 coarse launch 0.956 -> 0.852 ms, FP16 0.599 -> 0.586 ms; FP64 with
// pairs 0.97 ms against 0.94 ms in the aligned one-node form, so FP64 keeps that.
// Launches that also sum their writes for the flux controller (LatArgs::dsum) use
// DNSG_CS_PAIRED_DS_MASK: there FP64 pairs measured faster (the Re_tau=180 band
// launch 1.032 -> 0.937 ms, the coarse one 1.936 -> 1.898 ms).
#ifndef DNSG_CS_PAIRED_MASK
#define DNSG_CS_PAIRED_MASK 6
#endif
#ifndef DNSG_CS_PAIRED_DS_MASK
#define DNSG_CS_PAIRED_DS_MASK 7
#endif
__host__ __device__ __forceinline__ bool cs_paired(int nx, int elem, bool ds = false) {
    const int bit = elem == 8 ? 1 : (elem == 4 ? 2 : 4);
    return ((ds ? DNSG_CS_PAIRED_DS_MASK : DNSG_CS_PAIRED_MASK) & bit) != 0 && nx % (2 * CS_THREADS) == 0;
}
// collide-stream workgroups of one lattice with `rows` rows of nx nodes
__host__ __device__ __forceinline__ int cs_blocks(int nx, int rows, bool paired) {
    if (paired) return rows * (cs_chunks(nx) / 2);
    const int rpw = cs_rows_per_wg(nx);
    return rpw == 1 ? rows * cs_chunks(nx) : (rows + rpw - 1) / rpw;
}

struct Grid {
    void *buf[2] = {nullptr, nullptr};  // A/B lattices
    int cur = 0;                        // index of the current (pre-collision) lattice
    int nx = 0, ny = 0, nz = 0;
    size_t qs() const { return (size_t)nx * ny * nz; }
    size_t nodes() const { return qs(); }
    void *c() const { return buf[cur]; }
    void *n() const { return buf[cur ^ 1]; }
    void swap() { cur ^= 1; }
};

// Columns per tile of the two-substep band kernel (k_band2.hip).  When a step runs
// it, the bands' fused-reduction partials (LatArgs::red, ::dsum) are per (layer, row,
// B2_TX columns) instead of per CS_THREADS columns; dnsg_ctx::band_cols says which.
constexpr int B2_TX = 64;

// number of doubles in one diag record written by the reduction finalisers
constexpr int DIAG_WORDS = 9;
// per-row partial sums kept for one reduction pass
constexpr int MAX_PARTIAL_Q = 5;

}  // namespace dnsg

typedef struct dnsg_ctx dnsg_t;
struct dnsg_ctx {
    dnsg_params p;
    dnsg_dims d;
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false: the stream belongs to a queue serving many handles
    size_t elem = 8;  // bytes per stored distribution value

    dnsg::Grid coarse, lower, upper;
    // interface time levels [23][Yc][Xc], newest first, rotated by pointer swaps;
    // FP64 in the plain tier, FP32 (slot 0 = rho - 1) in the shifted tiers.  Slot 3
    // is a spare: the sampling of a step writes the spare, which the rotation makes
    // the newest level, so the three levels at the top of a step stay intact until
    // the step's wall-stress check has passed
    void *lev_lo[4] = {nullptr, nullptr, nullptr, nullptr};
    void *lev_up[4] = {nullptr, nullptr, nullptr, nullptr};
    // the steps use the two-substep band kernel (gratio 2); band_cols: columns per
    // fused-reduction partial in the bands' partials of the last launch that wrote them
    bool band2 = false;
    int band_cols = dnsg::CS_THREADS;
    bool cs_pair = true;  // FP32 / FP16: the paired collide-stream form where the rows allow it
    int8_t *wall_lo = nullptr, *wall_up = nullptr;     // [Yf][Xf]

    // reductions
    double *part = nullptr;     // row partials of the explicit reduction passes
    double *red = nullptr;      // sums written by the collide-stream kernels (RED_W per row chunk)
    double *seg = nullptr;      // first-stage segment sums of the flux reduction
    // flux controller from the writes of the previous step: destination-weighted
    // sums of the collide-stream launches [coarse | lower | upper] (2 per row chunk)
    // and of the injection (2 per workgroup); enc_next holds the next step's
    // G, <rho u>, <rho>, flowrate; flux_fused says d_force and enc_next belong to the
    // current state
    double *dsum = nullptr, *fsum = nullptr, *enc_next = nullptr;
    bool flux_fused = false;
    size_t part_cap = 0;        // doubles
    double *d_force = nullptr;  // device copy of V.force[0]
    double *d_diag = nullptr;   // DIAG_WORDS per step
    int diag_cap = 0;
    double *d_red = nullptr;    // small scratch for finalised vectors (statistics)
    double *h_pinned = nullptr; // pinned host staging for scalars
    size_t h_pinned_cap = 0;

    // running statistics accumulators [14][NZ] + count
    double *d_acc = nullptr;
    long acc_count = 0;
    // time sums of the budget profiles (dnsg_budget2d_accumulate): [19 + 3 + 30][NZ]
    std::vector<double> bud_acc;
    long bud_count = 0;

    // host copies of the band wall layers (never touched by the step) and their
    // mass contribution, for checkpoints and the mass check
    double *host_wall_lo = nullptr, *host_wall_up = nullptr;  // [Xf][Yf][19]
    double wall_mass = 0.;
    // filtered injection only (p.fltr): device copies of the two wall layers
    // [lower | upper][Xf][Yf][19], read when the filter's 5 layers reach a wall, and
    // f_neq of the 5 layers around each coincident layer [side][19][5][Yf][Xf]
    double *fwall = nullptr, *fneq = nullptr;

    // NaN halt: the wall-stress finaliser sets *d_halt when the stress is NaN, and while
    // halt_armed (inside dnsg_step / dnsg_run_timed*) every state-changing launch of
    // the later kernels is a no-op, so the state stays that of the failing step's top
    int *d_halt = nullptr;
    int *halt_flag = nullptr;  // the flag of the queued steps: d_halt, or the device's static scratch
    bool halt_armed = false;
    const int *halt() const { return halt_armed ? halt_flag : nullptr; }

    double force = 0.;  // host mirror of V.force[0]
    long ss = 0;        // next step index of the time loop
    bool have_state = false;
};

namespace dnsg {

// fused-reduction partials per (layer, row) of a band: in the current layout, and
// the most any layout needs (the buffers are sized for it)
inline int band_cpr(const dnsg_ctx *h) { return (h->d.Xf + h->band_cols - 1) / h->band_cols; }
inline int band_cpr_max(const dnsg_ctx *h) { return (h->d.Xf + B2_TX - 1) / B2_TX; }

// handle set-up and the per-step argument builders shared by the single-solve
// step and the serving queue
int create_ctx(const dnsg_params *p, hipStream_t ext, dnsg_t **out);
LatArgs coarse_args(dnsg_t *h);
void band_args(dnsg_t *h, LatArgs *L);
double *red_coarse(dnsg_t *h);
double *red_lower(dnsg_t *h);
double *red_upper(dnsg_t *h);
void rotate_levels(dnsg_t *h);

// launchers (stream-ordered, no host synchronisation)
hipError_t launch_collide_stream(hipStream_t s, const LatArgs *lats, int nlat, int prec, const int *halt = nullptr);
// both gratio-2 band substeps of one coarse step in one pass (k_band2.hip): bands[0]
// the lower band, bands[1] the upper band, as band_args builds them; the first fill
// must already be in their nxt lattices' interface layers; red / dsum in the B2_TX
// partial layout
hipError_t launch_band2(hipStream_t s, const LatArgs *bands, int prec, const int *halt = nullptr);
int band2_tiles(int nx, int ny, int prec);
hipError_t launch_init_lattice(hipStream_t s, dnsg_ctx *h, int grid, void *dst);
// gl = 1: the host array carries one ghost x-plane and y-row (the reference's in-memory
// layout); downloads write them as periodic images
hipError_t launch_aos_to_soa(hipStream_t s, const double *aos, void *soa, int nx, int ny, int nz_aos, int z0,
                             int nz, int prec, int gl = 0);
hipError_t launch_soa_to_aos(hipStream_t s, const void *soa, double *aos, int nx, int ny, int nz_aos, int z0,
                             int nz, int prec, int gl = 0);
hipError_t launch_sample(hipStream_t s, dnsg_ctx *h, const void *coarse_state, void *lo, void *up);
hipError_t launch_interface(hipStream_t s, dnsg_ctx *h, int iter);
// plain injection (ftcdtransfer.c) or, for a handle with p.fltr, the filtered one
// (ftcdtransfer_lesfltr.c); both write the same psum layout
hipError_t launch_ftcd(hipStream_t s, dnsg_ctx *h, double *psum = nullptr);
// the same kernels over several solves of one precision (one launch per
// BATCH_MAX items); the interface batch needs one gratio
hipError_t launch_sample_many(hipStream_t s, dnsg_ctx *const *hs, int n);
hipError_t launch_interface_many(hipStream_t s, dnsg_ctx *const *hs, int n, int iter);
hipError_t launch_ftcd_many(hipStream_t s, dnsg_ctx *const *hs, int n, double *psum = nullptr);
// the '-i' start: interpolation of a uniform-grid field [xdu][ydu][zdu+2][19] (device
// memory) onto the three state lattices, and the check that its stencils stay inside
int uniform_check(const dnsg_ctx *h, int xdu, int ydu, int zdu, std::string *err);
hipError_t launch_uniform(hipStream_t s, dnsg_ctx *h, const double *u, int xdu, int ydu, int zdu);
// velocity and density of grid g (0 coarse, 1 lower band, 2 upper band) as
// vel_print forms them, into [X + 2 pad][Y][Z][3] and [X + 2 pad][Y][Z] with pad
// periodic ghost planes in x (0: the vel-* / den-* file layout); needs fwall
hipError_t launch_snapshot(dnsg_ctx *h, int g, int pad, double *vel, double *den);
// device copies of the band wall layers (fwall), allocated and filled from the
// host copies on first use; every call that sets the state refreshes them
int ensure_fwall(dnsg_t *h);
// doubles of the filtered injection's f_neq scratch and of the wall-layer copies
size_t fltr_fneq_doubles(const dnsg_ctx *h);
size_t fltr_wall_doubles(const dnsg_ctx *h);
// workgroups of one injection launch (2 sides per handle), for the psum size
int ftcd_blocks(const dnsg_ctx *h);
// bulk momentum and density of each solve from its fused collide-stream sums:
// out[2 slot[i]] = <rho u_x>, out[2 slot[i] + 1] = <rho> (encalc2.c:109-131 weights)
hipError_t launch_flux_many(hipStream_t s, dnsg_ctx *const *hs, int n, double *out, const int *slot);
hipError_t launch_encalc2(hipStream_t s, dnsg_ctx *h, double *diag_rec);
hipError_t launch_wshear_fused(hipStream_t s, dnsg_ctx *h, double *diag_rec, const double *enc_rec = nullptr);
// constant force: encalc2 from the collide-stream kernels' row sums (RED_W per row) and
// wshearstr from the bands', in two launches with a shared first stage
hipError_t launch_reductions_fused(hipStream_t s, dnsg_ctx *h, double *diag_rec);
// G of the next step from the destination-weighted sums of this step's writes
hipError_t launch_flux_next(hipStream_t s, dnsg_ctx *h);
size_t dsum_doubles(const dnsg_ctx *h);
// encalc2 destination weights of one lattice (0 coarse, 1 lower band, 2 upper band)
void set_dest_weights(const dnsg_ctx *h, int grid, LatArgs *L);
hipError_t launch_wshear(hipStream_t s, dnsg_ctx *h, double *diag_rec);
hipError_t launch_mass(hipStream_t s, dnsg_ctx *h, double *out);
hipError_t launch_statistics(hipStream_t s, dnsg_ctx *h, double *out9nz);
hipError_t launch_fold_stats(hipStream_t s, double *acc, const double *st, int NZ);
size_t partial_doubles(const dnsg_ctx *h);
size_t enc_seg_doubles(const dnsg_ctx *h);

}  // namespace dnsg
