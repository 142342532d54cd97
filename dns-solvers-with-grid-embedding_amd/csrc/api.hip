// api.hip -- the dnsg_* C-ABI (include/dnsg.h): device state ownership, the
// kernel sequence of one coarse step, uploads/downloads in the reference's file
// layouts, and the reductions' host side.
//
// One handle = one GPU stream.  Every call is stream-ordered; calls that return
// host data synchronise the stream.  There is no CPU fallback: every compute
// path runs on the GPU or the call fails with DNSG_ERR_HIP.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "internal.hpp"

using namespace dnsg;

#define HIPCHK(h, expr)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                         \
            return e_ == hipErrorOutOfMemory ? DNSG_ERR_OOM : DNSG_ERR_HIP;                       \
        }                                                                                         \
    } while (0)

namespace {

constexpr int NVP = 23;

// parameter derivation, kept identical to main:139-155, 432-469
int derive(const dnsg_params *p, dnsg_dims *d, std::string *err) {
    memset(d, 0, sizeof(*d));
    auto fail = [&](const char *m) {
        if (err) *err = m;
        return DNSG_ERR_ARG;
    };
    if (p->gratio < 1 || p->gratio > 4) return fail("gratio must be 1..4");
    if (p->xdv < 2 || p->ydv < 2 || p->zdv < 5) return fail("xdv, ydv >= 2 and zdv >= 5 required");
    d->GR = p->gratio;
    d->Xc = p->xdv;
    d->Yc = p->ydv;
    d->zdc = p->zdv;
    d->Xf = d->GR * d->Xc;
    d->Yf = d->GR * d->Yc;
    d->zdf = p->zds + 1;
    if (d->zdf < 5 || d->zdf - 2 - d->GR < 0) return fail("zds must be >= max(4, gratio+1)");
    d->NZs = 2 * d->zdf + d->GR * (d->zdc - 4) + d->GR - 1;
    d->NZ = 2 * d->zdf + (d->zdc - 4);
    if (!(p->CFL > 0.)) return fail("CFL must be positive");
    const double ut = ((double)(0.5 * (d->NZs - 2.) / p->CFL)) / ((double)(d->GR));
    d->utime = (long)(int)ut;
    if (d->utime < 1) return fail("CFL gives fewer than one step per unit time");
    d->dt = d->GR * 1. / ((double)(0.5 * (d->NZs - 2.) / p->CFL));
    if (p->tau_c > 0.) {
        d->tau_c = p->tau_c;
        d->tau_f = d->GR * (p->tau_c - 0.5) + 0.5;
    } else {
        if (!(p->Rebs > 0.)) return fail("Rebs must be positive");
        d->tau_f = (3. / p->Rebs) * p->ub * (d->NZs - 2.) + 0.5;
        d->tau_c = (1. / ((double)d->GR)) * (3. / p->Rebs) * p->ub * (d->NZs - 2.) + 0.5;
    }
    if (!(d->tau_c > 0.5) || !(d->tau_f > 0.5)) return fail("relaxation times must exceed 1/2");
    d->omega_f = 1. / d->tau_f;
    d->omega_c = 1. / d->tau_c;
    const long nut = p->TEND - p->TSTR;
    d->up = (long)(nut * (double)d->utime);
    if (p->ITER > 0) d->up = p->ITER;
    d->n_coarse = (size_t)d->Xc * d->Yc * d->zdc * 19;
    d->n_band = (size_t)d->Xf * d->Yf * d->zdf * 19;
    d->n_plane = (size_t)3 * d->Xc * d->Yc * NVP;
    if (p->pattern == DNSG_PATTERN_RSLIP || p->pattern == DNSG_PATTERN_PSLIP) {
        const int P = p->gyplussize + p->wyplussize;
        if (p->gyplussize < 0 || p->wyplussize < 0 || P <= 0 || d->Yf % P) return fail("gyplussize+wyplussize must divide the fine spanwise extent");
    }
    if (p->pattern == DNSG_PATTERN_SSLIP || p->pattern == DNSG_PATTERN_PSLIP) {
        const int P = p->gxplussize + p->wxplussize;
        if (p->gxplussize < 0 || p->wxplussize < 0 || P <= 0 || d->Xf % P) return fail("gxplussize+wxplussize must divide the fine streamwise extent");
    }
    if (p->pattern < 0 || p->pattern > 3) return fail("unknown pattern");
    if (p->precision < 0 || p->precision > 2) return fail("unknown precision");
    if (p->force_mode < 0 || p->force_mode > 1) return fail("unknown force mode");
    if (p->fltr < 0 || p->fltr > 1) return fail("fltr must be 0 (NOFLTR) or 1 (FLTR)");
    // ftcdtransfer_lesfltr.c:22-28: the filter holds 5 fine layers in an array of GR+1 and
    // its weights sum to GR^3 only for GR = 4; its slab spans k = zdf-1-GR-2 .. GR+2
    if (p->fltr && d->GR != 4) return fail("FLTR requires gratio = 4");
    if (p->fltr && d->zdf - 3 - d->GR < 0) return fail("FLTR requires zds >= gratio + 2");
    return DNSG_OK;
}

// periodic stripe map of slip_ridges_init.c:37-42: solid (-1) except gas (-2)
// on [w/2 + n(g+w), w/2 + n(g+w) + g)
void stripe_map(int n, int gap, int width, std::vector<int> &g) {
    g.assign(n, -1);
    for (int j = width / 2; j < n; j += gap + width)
        for (int a = 0; a < gap && j + a < n; a++) g[j + a] = -2;
}

// wall codes of both walls, [Yf][Xf] (solid_init.c:46-64 + slip_*_init.c)
std::vector<int8_t> wall_codes(const dnsg_params *p, const dnsg_dims *d) {
    std::vector<int> sx(d->Xf, -1), sy(d->Yf, -1);
    if (p->pattern == DNSG_PATTERN_RSLIP) stripe_map(d->Yf, p->gyplussize, p->wyplussize, sy);
    if (p->pattern == DNSG_PATTERN_SSLIP) stripe_map(d->Xf, p->gxplussize, p->wxplussize, sx);
    if (p->pattern == DNSG_PATTERN_PSLIP) {
        stripe_map(d->Xf, p->gxplussize, p->wxplussize, sx);
        stripe_map(d->Yf, p->gyplussize, p->wyplussize, sy);
    }
    std::vector<int8_t> w((size_t)d->Xf * d->Yf);
    for (int j = 0; j < d->Yf; j++)
        for (int i = 0; i < d->Xf; i++) {
            int c = -1;
            if (p->pattern == DNSG_PATTERN_RSLIP) c = sy[j];
            else if (p->pattern == DNSG_PATTERN_SSLIP) c = sx[i];
            else if (p->pattern == DNSG_PATTERN_PSLIP) c = (sy[j] == -2) ? -2 : sx[i];
            w[(size_t)j * d->Xf + i] = (int8_t)c;
        }
    return w;
}

int ensure_diag(dnsg_t *h, int n) {
    if (n <= h->diag_cap) return DNSG_OK;
    if (h->d_diag) HIPCHK(h, hipFree(h->d_diag));
    h->d_diag = nullptr;
    h->diag_cap = 0;
    HIPCHK(h, hipMalloc(&h->d_diag, sizeof(double) * DIAG_WORDS * (size_t)n));
    h->diag_cap = n;
    return DNSG_OK;
}

int ensure_pinned(dnsg_t *h, size_t n) {
    if (n <= h->h_pinned_cap) return DNSG_OK;
    if (h->h_pinned) HIPCHK(h, hipHostFree(h->h_pinned));
    h->h_pinned = nullptr;
    h->h_pinned_cap = 0;
    HIPCHK(h, hipHostMalloc((void **)&h->h_pinned, sizeof(double) * n, hipHostMallocDefault));
    h->h_pinned_cap = n;
    return DNSG_OK;
}

}  // namespace

// ---- scratch of the allocation-free step ------------------------------------------
// dnsg_step_noalloc keeps the NaN halt flag (first 256 bytes) and the records of up to
// DNSG_NOALLOC_MAX_STEPS queued steps in this file-scope device array.  Each GPU holds
// its own copy until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host that name is a stand-in, not a device address): the
// address on each device comes from hipGetSymbolAddress, taken once per device by
// dnsg_create with that device current, and kept in scratch_addr.
namespace {
constexpr size_t NOALLOC_REC_OFS = 32;  // doubles before the records: the halt flag, padded to 256 B
constexpr size_t NOALLOC_WORDS = NOALLOC_REC_OFS + (size_t)DNSG_NOALLOC_MAX_STEPS * DIAG_WORDS;
constexpr int NOALLOC_DEVICES = 64;
}  // namespace
static __device__ __align__(256) double scratch[NOALLOC_WORDS];
static double *scratch_addr[NOALLOC_DEVICES];
static hipError_t scratch_status[NOALLOC_DEVICES];
static std::once_flag scratch_once[NOALLOC_DEVICES];

// the device address of this handle's device's scratch (h->device is current when
// dnsg_create calls it); base may be null
static int noalloc_scratch(dnsg_t *h, double **base) {
    if (h->device < 0 || h->device >= NOALLOC_DEVICES) {
        h->err = "device index beyond the scratch table";
        return DNSG_ERR_ARG;
    }
    std::call_once(scratch_once[h->device], [h]() {
        void *p = nullptr;
        scratch_status[h->device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        scratch_addr[h->device] = static_cast<double *>(p);
    });
    if (scratch_status[h->device] != hipSuccess || !scratch_addr[h->device]) {
        h->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(scratch_status[h->device]);
        return DNSG_ERR_HIP;
    }
    if (base) *base = scratch_addr[h->device];
    return DNSG_OK;
}

namespace dnsg {

LatArgs coarse_args(dnsg_t *h) {
    LatArgs L;
    L.cur = h->coarse.c();
    L.nxt = h->coarse.n();
    L.nx = h->coarse.nx; L.ny = h->coarse.ny; L.nz = h->coarse.nz;
    L.lo_kind = BOUND_DROP; L.hi_kind = BOUND_DROP;
    L.wlo = nullptr; L.whi = nullptr;
    L.gscale = 1.;
    L.omega = h->d.omega_c;
    L.skip_lo = 1;  // coarse k = 0, zdc-1 are rewritten by the fine->coarse injection
    L.skip_hi = 1;
    L.red = nullptr;
    L.force = h->d_force;
    L.dsum = nullptr;
    L.paired = h->cs_pair ? (cs_paired(L.nx, h->elem) ? 1 : 0) | (cs_paired(L.nx, h->elem, true) ? 2 : 0) : 0;
    set_dest_weights(h, 0, &L);
    return L;
}

void band_args(dnsg_t *h, LatArgs *L) {
    const double gs = 1. / ((double)h->d.GR);  // xfcomputations' GR argument (main:1349)
    for (int b = 0; b < 2; b++) {
        Grid &g = b == 0 ? h->lower : h->upper;
        L[b].cur = g.c();
        L[b].nxt = g.n();
        L[b].nx = g.nx; L[b].ny = g.ny; L[b].nz = g.nz;
        L[b].lo_kind = b == 0 ? BOUND_WALL : BOUND_DROP;
        L[b].hi_kind = b == 0 ? BOUND_DROP : BOUND_WALL;
        L[b].wlo = b == 0 ? h->wall_lo : nullptr;
        L[b].whi = b == 0 ? nullptr : h->wall_up;
        L[b].gscale = gs;
        L[b].omega = h->d.omega_f;
        // the interface layer (lower: top stored layer, upper: layer 0) is rewritten
        // by the coarse->fine transfer after every substep
        L[b].skip_lo = b == 0 ? 0 : 1;
        L[b].skip_hi = b == 0 ? 1 : 0;
        L[b].red = nullptr;
        L[b].force = h->d_force;
        L[b].dsum = nullptr;
        L[b].paired = h->cs_pair ? (cs_paired(g.nx, h->elem) ? 1 : 0) | (cs_paired(g.nx, h->elem, true) ? 2 : 0) : 0;
        set_dest_weights(h, 1 + b, &L[b]);
    }
}

double *red_coarse(dnsg_t *h) { return h->red; }
double *red_lower(dnsg_t *h) { return h->red + (size_t)h->coarse.ny * h->coarse.nz * cs_chunks(h->coarse.nx) * RED_W; }
double *red_upper(dnsg_t *h) { return red_lower(h) + (size_t)h->lower.ny * h->lower.nz * band_cpr(h) * RED_W; }

void rotate_levels(dnsg_t *h) {  // xcomputations.c:548-556
    // the spare (slot 3) becomes the newest level, the oldest the spare
    void *t = h->lev_lo[3];
    h->lev_lo[3] = h->lev_lo[2]; h->lev_lo[2] = h->lev_lo[1]; h->lev_lo[1] = h->lev_lo[0]; h->lev_lo[0] = t;
    t = h->lev_up[3];
    h->lev_up[3] = h->lev_up[2]; h->lev_up[2] = h->lev_up[1]; h->lev_up[1] = h->lev_up[0]; h->lev_up[0] = t;
}

}  // namespace dnsg

namespace {

int do_xcomputations(dnsg_t *h) {
    LatArgs L = coarse_args(h);
    HIPCHK(h, launch_collide_stream(h->stream, &L, 1, h->p.precision));
    rotate_levels(h);
    HIPCHK(h, launch_sample(h->stream, h, h->coarse.n(), h->lev_lo[0], h->lev_up[0]));
    return DNSG_OK;
}

int do_xfcomputations(dnsg_t *h) {
    LatArgs L[2];
    band_args(h, L);
    HIPCHK(h, launch_collide_stream(h->stream, L, 2, h->p.precision));
    return DNSG_OK;
}

// optional section timing: events recorded on the handle's stream between the
// kernel groups of one step
enum { SEC_REDUCE = 0, SEC_COARSE = 1, SEC_SAMPLE = 2, SEC_BAND = 3, SEC_IFACE = 4, SEC_N = 5 };
// Every group (only < 0): one event at the start of the step and one after each
// group, so the groups are timed back to back.  One group (only = k): an event
// just before and just after each launch of group k, and no others.
struct SecTimer {
    std::vector<hipEvent_t> ev;
    std::vector<int> tag;  // section that ends at ev[i]; -1 starts an interval
    int only = -1;
    explicit SecTimer(int group = -1) : only(group) {}
    SecTimer(const SecTimer &) = delete;
    SecTimer &operator=(const SecTimer &) = delete;
    ~SecTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void record(hipStream_t s, int t) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, s);
        ev.push_back(e);
        tag.push_back(t);
    }
    void start_step(hipStream_t s) {
        if (only < 0) record(s, -1);
    }
    void begin(hipStream_t s, int t) {
        if (t == only) record(s, -1);
    }
    void end(hipStream_t s, int t) {
        if (only < 0 || t == only) record(s, t);
    }
};

// average ms per launch (group) of each section over the recorded steps; the
// stream must have passed the last event
static int sec_average(dnsg_t *h, const SecTimer &T, double *ms) {
    double acc[SEC_N] = {0., 0., 0., 0., 0.};
    int cnt[SEC_N] = {0, 0, 0, 0, 0};
    for (size_t i = 1; i < T.ev.size(); i++) {
        if (T.tag[i] < 0) continue;
        float t = 0.f;
        const hipError_t e = hipEventElapsedTime(&t, T.ev[i - 1], T.ev[i]);
        if (e != hipSuccess) {
            h->err = std::string("event timing: ") + hipGetErrorString(e);
            return DNSG_ERR_HIP;
        }
        acc[T.tag[i]] += t;
        cnt[T.tag[i]]++;
    }
    for (int k = 0; k < SEC_N; k++) ms[k] = cnt[k] ? acc[k] / cnt[k] : 0.;
    return DNSG_OK;
}

// One coarse step of main:1256-1375; rec receives the diag record on the device.
// The reductions of encalc2 and wshearstr read the state at the top of the step,
// which is exactly what the coarse collide-stream and the first band substep
// load: they sum it there (per-row partials) instead of re-reading the lattices.
// Only the flux controller needs G = ub - <rho u> before the coarse collision, so
// in that mode encalc2 keeps its own reduction pass.
// The coarse collide-stream and the first band substep touch disjoint lattices
// and read the same force, so they run as one launch over the three lattices
// (xcomputations and the first xfcomputations of the reference's loop).
// One coarse step with the two-substep band kernel (gratio 2, k_band2.hip): the
// coarse collide-stream alone, the sampling and the first fill (into the bands'
// output lattices), both band substeps in one pass, then the reductions (the bands'
// sums come from that pass), the second fill and the injection.  Every launch before
// the wall-stress check writes only the lattices' other buffers and the spare
// interface level, so a NaN there leaves the state of the step's top intact.
int do_step_band2(dnsg_t *h, double *rec, SecTimer *T) {
    const bool flux = h->p.force_mode == DNSG_FORCE_FLUX;
    const bool from_writes = flux && h->flux_fused;
    h->band_cols = B2_TX;
    if (T) T->start_step(h->stream);
    if (flux && !from_writes) {
        if (T) T->begin(h->stream, SEC_REDUCE);
        HIPCHK(h, launch_encalc2(h->stream, h, rec));
        if (T) T->end(h->stream, SEC_REDUCE);
    }
    LatArgs C = coarse_args(h);
    if (!flux) C.red = red_coarse(h);
    if (flux) C.dsum = h->dsum;
    if (T) T->begin(h->stream, SEC_COARSE);
    HIPCHK(h, launch_collide_stream(h->stream, &C, 1, h->p.precision, h->halt()));
    if (T) T->end(h->stream, SEC_COARSE);
    rotate_levels(h);
    if (T) T->begin(h->stream, SEC_SAMPLE);
    HIPCHK(h, launch_sample(h->stream, h, h->coarse.n(), h->lev_lo[0], h->lev_up[0]));
    if (T) T->end(h->stream, SEC_SAMPLE);
    if (T) T->begin(h->stream, SEC_IFACE);
    HIPCHK(h, launch_interface(h->stream, h, 1));
    if (T) T->end(h->stream, SEC_IFACE);
    LatArgs B[2];
    band_args(h, B);
    B[0].red = red_lower(h);
    B[1].red = red_upper(h);
    if (flux) {
        B[0].dsum = h->dsum + 2 * (size_t)h->d.Yc * h->d.zdc * cs_chunks(h->d.Xc);
        B[1].dsum = B[0].dsum + 2 * (size_t)h->d.Yf * (h->d.zdf - 1) * band_cpr(h);
    }
    if (T) T->begin(h->stream, SEC_BAND);
    HIPCHK(h, launch_band2(h->stream, B, h->p.precision, h->halt()));
    if (T) T->end(h->stream, SEC_BAND);
    if (T) T->begin(h->stream, SEC_REDUCE);
    if (!flux) HIPCHK(h, launch_reductions_fused(h->stream, h, rec));
    else HIPCHK(h, launch_wshear_fused(h->stream, h, rec, from_writes ? h->enc_next : nullptr));
    if (T) T->end(h->stream, SEC_REDUCE);
    if (T) T->begin(h->stream, SEC_IFACE);
    HIPCHK(h, launch_interface(h->stream, h, 2));
    HIPCHK(h, launch_ftcd(h->stream, h, flux ? h->fsum : nullptr));
    if (T) T->end(h->stream, SEC_IFACE);
    if (flux) {
        if (T) T->begin(h->stream, SEC_REDUCE);
        HIPCHK(h, launch_flux_next(h->stream, h));
        if (T) T->end(h->stream, SEC_REDUCE);
    }
    h->lower.swap();
    h->upper.swap();
    h->flux_fused = flux;
    h->coarse.swap();
    h->ss++;
    return DNSG_OK;
}

int do_step(dnsg_t *h, double *rec, SecTimer *T = nullptr) {
    if (h->band2) return do_step_band2(h, rec, T);
    const bool flux = h->p.force_mode == DNSG_FORCE_FLUX;
    // flux mode: G for this step comes from the previous step's writes when that
    // step ran here (flux_fused), else from a pass over the lattices (first step,
    // or after the state was changed from outside)
    const bool from_writes = flux && h->flux_fused;
    const int GR = h->d.GR;
    h->band_cols = CS_THREADS;
    if (T) T->start_step(h->stream);
    if (flux && !from_writes) {
        if (T) T->begin(h->stream, SEC_REDUCE);
        HIPCHK(h, launch_encalc2(h->stream, h, rec));
        if (T) T->end(h->stream, SEC_REDUCE);
    }
    double *dl = h->dsum + 2 * (size_t)h->d.Yc * h->d.zdc * cs_chunks(h->d.Xc);
    double *du = dl + 2 * (size_t)h->d.Yf * (h->d.zdf - 1) * band_cpr(h);
    for (int iter = 1; iter <= GR; iter++) {
        LatArgs L[3];
        LatArgs *B = L;  // the band pair of this substep
        int nl = 2;
        if (iter == 1) {
            L[0] = coarse_args(h);
            B = L + 1;
            nl = 3;
            if (!flux) L[0].red = red_coarse(h);
            if (flux) L[0].dsum = h->dsum;
        }
        band_args(h, B);
        if (iter == 1) {
            B[0].red = red_lower(h);
            B[1].red = red_upper(h);
        }
        if (flux && iter == GR) {  // the last substep writes the bands' next state
            B[0].dsum = dl;
            B[1].dsum = du;
        }
        if (T) T->begin(h->stream, iter == 1 ? SEC_COARSE : SEC_BAND);
        HIPCHK(h, launch_collide_stream(h->stream, L, nl, h->p.precision, h->halt()));
        if (T) T->end(h->stream, iter == 1 ? SEC_COARSE : SEC_BAND);
        if (iter == 1) {
            // the reductions first: a NaN wall stress halts every later launch, the
            // sampling included, so the state stays that at the top of this step
            if (T) T->begin(h->stream, SEC_REDUCE);
            if (!flux) HIPCHK(h, launch_reductions_fused(h->stream, h, rec));
            else HIPCHK(h, launch_wshear_fused(h->stream, h, rec, from_writes ? h->enc_next : nullptr));
            if (T) T->end(h->stream, SEC_REDUCE);
            rotate_levels(h);
            if (T) T->begin(h->stream, SEC_SAMPLE);
            HIPCHK(h, launch_sample(h->stream, h, h->coarse.n(), h->lev_lo[0], h->lev_up[0]));
            if (T) T->end(h->stream, SEC_SAMPLE);
        }
        if (T) T->begin(h->stream, SEC_IFACE);
        HIPCHK(h, launch_interface(h->stream, h, iter));
        if (iter == GR) HIPCHK(h, launch_ftcd(h->stream, h, flux ? h->fsum : nullptr));
        if (T) T->end(h->stream, SEC_IFACE);
        h->lower.swap();
        h->upper.swap();
    }
    if (flux) {
        if (T) T->begin(h->stream, SEC_REDUCE);
        HIPCHK(h, launch_flux_next(h->stream, h));
        if (T) T->end(h->stream, SEC_REDUCE);
    }
    h->flux_fused = flux;
    h->coarse.swap();
    h->ss++;
    return DNSG_OK;
}

void fill_diag(const dnsg_t *h, const double *rec, long ss, dnsg_diag *d) {
    d->time = ((double)h->p.TSTR + h->d.dt * ss);
    d->force = rec[1];
    d->momentum = rec[2];
    d->density = rec[3];
    d->flowrate = rec[4];
    d->str_up = rec[5];
    d->str_down = rec[6];
    d->stress = rec[7];
    d->pgrad = rec[8];
}

// The pointer state at the top of a step: which buffer of each grid is current,
// the interface levels, and whether the flux controller's G for it is formed.
struct StepTop {
    int cur[3];
    void *lo[4], *up[4];
    bool flux_fused;
};

StepTop step_top(const dnsg_t *h) {
    StepTop t;
    t.cur[0] = h->coarse.cur; t.cur[1] = h->lower.cur; t.cur[2] = h->upper.cur;
    for (int l = 0; l < 4; l++) { t.lo[l] = h->lev_lo[l]; t.up[l] = h->lev_up[l]; }
    t.flux_fused = h->flux_fused;
    return t;
}

// the top of the step s steps after top t: do_step swaps the coarse buffers once and
// the band buffers GR times (once with the two-substep band kernel), rotates the
// interface levels once, and leaves the flux controller's G formed from its writes
// in flux mode
StepTop top_after(const dnsg_t *h, StepTop t, long s) {
    const int band_swaps = h->band2 ? 1 : h->d.GR;
    for (long k = 0; k < s; k++) {
        t.cur[0] ^= 1;
        if (band_swaps & 1) { t.cur[1] ^= 1; t.cur[2] ^= 1; }
        void *a = t.lo[3]; t.lo[3] = t.lo[2]; t.lo[2] = t.lo[1]; t.lo[1] = t.lo[0]; t.lo[0] = a;
        a = t.up[3]; t.up[3] = t.up[2]; t.up[2] = t.up[1]; t.up[1] = t.up[0]; t.up[0] = a;
        t.flux_fused = h->p.force_mode == DNSG_FORCE_FLUX;
    }
    return t;
}

void restore_top(dnsg_t *h, const StepTop &t) {
    h->coarse.cur = t.cur[0]; h->lower.cur = t.cur[1]; h->upper.cur = t.cur[2];
    for (int l = 0; l < 4; l++) { h->lev_lo[l] = t.lo[l]; h->lev_up[l] = t.up[l]; }
    h->flux_fused = t.flux_fused;
}

// clears the halt flag with a kernel rather than a memset node, so a caller that
// captures the queued steps into a graph replays the clear too
__global__ void k_clear_flag(int *f) {
    if (threadIdx.x == 0) f[0] = 0;
}

// Where the queued steps keep their records and halt flag.  dnsg_step and the timed
// runs use the handle's own buffers (grown on demand) and pinned staging;
// dnsg_step_noalloc passes the device's static scratch array and copies the records
// back through a fixed host buffer on the stack.
struct StepBufs {
    double *rec;  // device, DIAG_WORDS per queued step
    int *halt;    // device halt flag
    bool pinned;  // records come back through h->h_pinned (else in chunks through the stack)
};

// Queue m steps without host synchronisation, then copy the m diag records back.
// A NaN wall stress in step s (wshearstr.c:139-143 aborts there) sets the device
// halt flag, after which every state-changing launch of the queued steps returns
// at once; the host then rewinds its buffer and level pointers to the top of step
// s, so the state, ss and force are those the reference holds when it aborts, and
// the records up to and including step s are filled.  With ms, the HIP-event time
// of the m steps; with T, section events inside them.  bufs = null: the handle's
// own buffers.
int run_steps(dnsg_t *h, int m, dnsg_diag *diag, float *ms, SecTimer *T, const StepBufs *bufs = nullptr) {
    int rc;
    StepBufs B;
    if (bufs) {
        B = *bufs;
    } else {
        if ((rc = ensure_diag(h, m))) return rc;
        if ((rc = ensure_pinned(h, (size_t)m * DIAG_WORDS))) return rc;
        B = {h->d_diag, h->d_halt, true};
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIPCHK(h, hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            h->err = "hipEventCreate failed";
            return DNSG_ERR_HIP;
        }
    }
    const StepTop top0 = step_top(h);
    const long ss0 = h->ss;
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, h->stream, B.halt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && ms) e = hipEventRecord(e0, h->stream);
    rc = e == hipSuccess ? DNSG_OK : DNSG_ERR_HIP;
    if (rc) h->err = std::string("step set-up: ") + hipGetErrorString(e);
    h->halt_flag = B.halt;
    h->halt_armed = true;
    for (int s = 0; s < m && !rc; s++) rc = do_step(h, B.rec + (size_t)s * DIAG_WORDS, T);
    h->halt_armed = false;
    h->halt_flag = h->d_halt;
    if (!rc && ms) {
        e = hipEventRecord(e1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
        if (e != hipSuccess) {
            h->err = std::string("step timing: ") + hipGetErrorString(e);
            rc = DNSG_ERR_HIP;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    constexpr int CH = 64;  // records per copy through the stack buffer
    double stage[CH * DIAG_WORDS];
    double last_force = h->force;
    for (int c0 = 0; c0 < m; c0 += CH) {
        const int nc = B.pinned ? m : (m - c0 < CH ? m - c0 : CH);
        double *hb = B.pinned ? h->h_pinned : stage;
        HIPCHK(h, hipMemcpyAsync(hb, B.rec + (size_t)c0 * DIAG_WORDS, sizeof(double) * DIAG_WORDS * nc,
                                 hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (int i = 0; i < nc; i++) {
            const int s = c0 + i;
            const double *rec = hb + (size_t)i * DIAG_WORDS;
            if (diag) fill_diag(h, rec, ss0 + s, diag + s);
            last_force = rec[1];
            if (isnan(rec[7])) {
                restore_top(h, top_after(h, top0, s));
                h->ss = ss0 + s;
                h->force = rec[1];
                char buf[96];
                snprintf(buf, sizeof buf, "NaN in wall shear stress at step %ld", ss0 + s);
                h->err = buf;
                return DNSG_ERR_NAN;
            }
        }
        if (B.pinned) break;
    }
    h->force = last_force;
    return DNSG_OK;
}

// Host layouts of the state.  File layout (ghost = false): lattices [X][Y][Z][19],
// interface levels [Xc][Yc][23].  The reference's in-memory layout (ghost = true,
// main:145-151, 473-883): lattices [X+1][Y+1][Z][19] whose last x-plane and y-row
// are ghost cells (the halo of the periodic neighbour, main:181-189), interface
// levels [Xc+6][Yc+6][23] with the interior at offset (2, 2) (Fbi,
// definitions.h:132; V.lin allocation main:827-833).  Uploads read the interior;
// downloads also write every ghost cell as its periodic image.
struct HostLayout {
    bool ghost = false;
    int gl() const { return ghost ? 1 : 0; }              // extra lattice planes in x and y
    size_t lev_at(int i, int j, int v, int Xc, int Yc) const {  // interior (i, j) of a level
        return ghost ? ((size_t)(i + 2) * (Yc + 6) + (j + 2)) * 23 + v : ((size_t)i * Yc + j) * 23 + v;
    }
};

// three host levels (newest first) <-> three device planes [23][Yc][Xc]; the shifted
// tiers keep FP32 planes with rho - 1 in slot 0
int put_levels(dnsg_t *h, const double *const *lv, void *const *dev, HostLayout L = {}) {
    const int Xc = h->d.Xc, Yc = h->d.Yc;
    const size_t np = (size_t)Xc * Yc * NVP;
    const bool f32 = h->p.precision != DNSG_PREC_FP64;
    std::vector<double> t(np);
    std::vector<float> tf(f32 ? np : 0);
    for (int l = 0; l < 3; l++) {
        const double *src = lv[l];
        for (int i = 0; i < Xc; i++)
            for (int j = 0; j < Yc; j++)
                for (int v = 0; v < NVP; v++) {
                    const double x = src[L.lev_at(i, j, v, Xc, Yc)];
                    const size_t d = ((size_t)v * Yc + j) * Xc + i;
                    if (f32) tf[d] = (float)(v == 0 ? x - 1. : x);
                    else t[d] = x;
                }
        if (f32) HIPCHK(h, hipMemcpyAsync(dev[l], tf.data(), np * 4, hipMemcpyHostToDevice, h->stream));
        else HIPCHK(h, hipMemcpyAsync(dev[l], t.data(), np * 8, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return DNSG_OK;
}

int get_levels(dnsg_t *h, double *const *lv, void *const *dev, HostLayout L = {}) {
    const int Xc = h->d.Xc, Yc = h->d.Yc;
    const size_t np = (size_t)Xc * Yc * NVP;
    const bool f32 = h->p.precision != DNSG_PREC_FP64;
    std::vector<double> t(np);
    std::vector<float> tf(f32 ? np : 0);
    for (int l = 0; l < 3; l++) {
        if (f32) HIPCHK(h, hipMemcpyAsync(tf.data(), dev[l], np * 4, hipMemcpyDeviceToHost, h->stream));
        else HIPCHK(h, hipMemcpyAsync(t.data(), dev[l], np * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        double *dst = lv[l];
        const int g = L.ghost ? 2 : 0, ge = L.ghost ? 4 : 0;  // halo cells before / after the interior
        for (int i = -g; i < Xc + ge; i++)
            for (int j = -g; j < Yc + ge; j++) {
                const int ii = (i + Xc) % Xc, jj = (j + Yc) % Yc;  // periodic image
                for (int v = 0; v < NVP; v++) {
                    const size_t d = ((size_t)v * Yc + jj) * Xc + ii;
                    const size_t at = L.ghost ? ((size_t)(i + 2) * (Yc + 6) + (j + 2)) * 23 + v
                                              : ((size_t)i * Yc + j) * 23 + v;
                    dst[at] = f32 ? (v == 0 ? 1. + (double)tf[d] : (double)tf[d]) : t[d];
                }
            }
    }
    return DNSG_OK;
}

// AoS [X(+1)][Y(+1)][zaos][19] host array <-> SoA device lattice (stored layers z0..z0+nz-1)
int put_lattice(dnsg_t *h, const double *aos, Grid &g, int zaos, int z0, HostLayout L = {}) {
    const int gl = L.gl();
    const size_t n = (size_t)(g.nx + gl) * (g.ny + gl) * zaos * 19;
    double *stage = nullptr;
    HIPCHK(h, hipMalloc(&stage, n * 8));
    hipError_t e = hipMemcpyAsync(stage, aos, n * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = launch_aos_to_soa(h->stream, stage, g.c(), g.nx, g.ny, zaos, z0, g.nz, h->p.precision, gl);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(stage);
    HIPCHK(h, e);
    return DNSG_OK;
}

int get_lattice(dnsg_t *h, double *aos, const Grid &g, int zaos, int z0, HostLayout L = {}) {
    const int gl = L.gl();
    const size_t n = (size_t)(g.nx + gl) * (g.ny + gl) * zaos * 19;
    double *stage = nullptr;
    HIPCHK(h, hipMalloc(&stage, n * 8));
    hipError_t e = hipMemsetAsync(stage, 0, n * 8, h->stream);
    if (e == hipSuccess)
        e = launch_soa_to_aos(h->stream, g.c(), stage, g.nx, g.ny, zaos, z0, g.nz, h->p.precision, gl);
    if (e == hipSuccess) e = hipMemcpyAsync(aos, stage, n * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(stage);
    HIPCHK(h, e);
    return DNSG_OK;
}

// keep the wall layer k of a band AoS array on the host; returns its mass
double keep_wall(const dnsg_t *h, const double *aos, int k, double *dst, HostLayout L = {}) {
    const int Xf = h->d.Xf, Yf = h->d.Yf, zdf = h->d.zdf, ldy = Yf + L.gl();
    double m = 0.;
    for (int i = 0; i < Xf; i++)
        for (int j = 0; j < Yf; j++) {
            const double *src = aos + (((size_t)i * ldy + j) * zdf + k) * 19;
            double r = 0.;
            for (int a = 0; a < 19; a++) {
                dst[((size_t)i * Yf + j) * 19 + a] = src[a];
                r += src[a];
            }
            m += r;
        }
    return m;
}

void place_wall(const dnsg_t *h, double *aos, int k, const double *src, HostLayout L = {}) {
    const int Xf = h->d.Xf, Yf = h->d.Yf, zdf = h->d.zdf, gl = L.gl(), ldy = Yf + gl;
    for (int i = 0; i < Xf + gl; i++)
        for (int j = 0; j < Yf + gl; j++)
            memcpy(aos + (((size_t)i * ldy + j) * zdf + k) * 19, src + ((size_t)(i % Xf) * Yf + (j % Yf)) * 19, 19 * 8);
}

}  // namespace

extern "C" {

int dnsg_params_default(dnsg_params *p) {
    if (!p) return DNSG_ERR_ARG;
    memset(p, 0, sizeof(*p));
    // definitions.h:34-62 with the grid of the 64x32x32 channel on one GPU
    p->Rebs = 7200.000;
    p->ub = 0.0866034;
    p->zdv = 32;
    p->ydv = 32;
    p->xdv = 64;
    p->TSTR = 0;
    p->TEND = 1;
    p->CFL = 0.1;
    p->gratio = 2;
    p->zds = 8;
    p->gyplussize = 56;
    p->wyplussize = 8;
    p->gxplussize = 4;
    p->wxplussize = 4;
    p->pattern = DNSG_PATTERN_NONE;
    p->ITER = 0;
    p->force_mode = DNSG_FORCE_FLUX;
    p->g_const = 0.;
    p->tau_c = 0.;
    p->precision = DNSG_PREC_FP64;
    p->init_kind = DNSG_INIT_REST;
    p->seed = 1234;
    p->u_max = 0.05;
    p->noise = 0.05;
    return DNSG_OK;
}

int dnsg_dims_from_params(const dnsg_params *p, dnsg_dims *d) {
    if (!p || !d) return DNSG_ERR_ARG;
    return derive(p, d, nullptr);
}

const char *dnsg_last_error(const dnsg_t *h) { return h ? h->err.c_str() : "null handle"; }

int dnsg_get_dims(const dnsg_t *h, dnsg_dims *d) {
    if (!h || !d) return DNSG_ERR_ARG;
    *d = h->d;
    return DNSG_OK;
}

// The node codes as the reference holds them (solid_init.c:14-19: -1 no-slip wall,
// -2 slip wall, 0 fluid, 1 periodic / bounded), read from what the kernels of this
// handle use: the wall codes from the device wall maps, the bounded layers from the
// collide-stream's boundary set-up (the layers whose outgoing populations are dropped
// and rewritten by an interface transfer), and 1 on the periodic x/y edge planes of
// the one-rank ghosted extents.
int dnsg_node_codes(dnsg_t *h, int *s_c, int *s_l, int *s_u) {
    if (!h || !h->stream) return DNSG_ERR_ARG;
    const int Xc = h->d.Xc, Yc = h->d.Yc, Xf = h->d.Xf, Yf = h->d.Yf, zdc = h->d.zdc, zdf = h->d.zdf;
    std::vector<int8_t> wl((size_t)Xf * Yf), wu((size_t)Xf * Yf);
    LatArgs C = coarse_args(h), B[2];
    band_args(h, B);
    if (B[0].lo_kind != BOUND_WALL || B[1].hi_kind != BOUND_WALL || B[0].wlo == nullptr || B[1].whi == nullptr) {
        h->err = "band wall boundaries not set up";
        return DNSG_ERR_STATE;
    }
    HIPCHK(h, hipMemcpyAsync(wl.data(), B[0].wlo, wl.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(wu.data(), B[1].whi, wu.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int xdc = Xc + 1, ydc = Yc + 1, xdf = Xf + 1, ydf = Yf + 1;
    if (s_c)
        for (int i = 0; i < xdc; i++)
            for (int j = 0; j < ydc; j++)
                for (int k = 0; k < zdc; k++) {
                    const bool edge = i == 0 || i == xdc - 1 || j == 0 || j == ydc - 1;
                    const bool bounded = (k == 0 && C.skip_lo) || (k == zdc - 1 && C.skip_hi);
                    s_c[((size_t)i * ydc + j) * zdc + k] = (edge || bounded) ? 1 : 0;
                }
    // lower band: stored layer m = k - 1 (k = 0 is the wall); upper band: m = k
    for (int b = 0; b < 2; b++) {
        int *s = b == 0 ? s_l : s_u;
        if (!s) continue;
        const LatArgs &L = B[b];
        const std::vector<int8_t> &w = b == 0 ? wl : wu;
        const int kwall = b == 0 ? 0 : zdf - 1, kofs = b == 0 ? 1 : 0;
        for (int i = 0; i < xdf; i++)
            for (int j = 0; j < ydf; j++)
                for (int k = 0; k < zdf; k++) {
                    int c;
                    if (k == kwall) {
                        c = w[(size_t)(j % Yf) * Xf + (i % Xf)];  // the ghost planes hold the periodic image
                    } else {
                        const int m = k - kofs;
                        const bool edge = i == 0 || i == xdf - 1 || j == 0 || j == ydf - 1;
                        const bool bounded = (m == 0 && L.skip_lo) || (m == L.nz - 1 && L.skip_hi);
                        c = (edge || bounded) ? 1 : 0;
                    }
                    s[((size_t)i * ydf + j) * zdf + k] = c;
                }
    }
    return DNSG_OK;
}

int dnsg_destroy(dnsg_t *h) {
    if (!h) return DNSG_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    dnsg::Grid *gs[3] = {&h->coarse, &h->lower, &h->upper};
    for (auto *g : gs)
        for (int b = 0; b < 2; b++)
            if (g->buf[b]) (void)hipFree(g->buf[b]);
    for (int l = 0; l < 4; l++) {
        if (h->lev_lo[l]) (void)hipFree(h->lev_lo[l]);
        if (h->lev_up[l]) (void)hipFree(h->lev_up[l]);
    }
    void *ptrs[] = {h->wall_lo, h->wall_up, h->part, h->red, h->seg, h->dsum, h->fsum, h->enc_next,
                    h->d_force, h->d_halt, h->d_diag, h->d_red, h->d_acc, h->fwall, h->fneq};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (h->h_pinned) (void)hipHostFree(h->h_pinned);
    free(h->host_wall_lo);
    free(h->host_wall_up);
    if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return DNSG_OK;
}

int dnsg_create(const dnsg_params *p, dnsg_t **out) { return dnsg::create_ctx(p, nullptr, out); }

}  // extern "C"

// a handle on the caller's stream (ext) or on a stream of its own (ext = null)
int dnsg::create_ctx(const dnsg_params *p, hipStream_t ext, dnsg_t **out) {
    if (!p || !out) return DNSG_ERR_ARG;
    *out = nullptr;
    dnsg_t *h = new (std::nothrow) dnsg_ctx();
    if (!h) return DNSG_ERR_OOM;
    h->p = *p;
    int rc = derive(p, &h->d, &h->err);
    if (rc) {
        *out = h;  // keep the handle so the caller can read the message
        return rc;
    }
    *out = h;
    HIPCHK(h, hipGetDevice(&h->device));
    if (ext) {
        h->stream = ext;
        h->own_stream = false;
    } else {
        HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    }
    h->elem = p->precision == DNSG_PREC_FP64 ? 8 : (p->precision == DNSG_PREC_FP32 ? 4 : 2);
    {
        // gratio 2: both band substeps in one pass (k_band2.hip), by default in the FP64
        // and FP32 constant-force configurations, where it is the faster path (FP16
        // storage and the flux controller: one launch per substep); DNSG_BAND2=1 /
        // DNSG_BAND2=0 in the environment selects it / the one-launch-per-substep path
        // for any configuration (the same state; the reductions' partials differ)
        // FP32 / FP16 rows of a multiple of 2 CS_THREADS nodes: the paired collide-stream
        // form (k_stream.hip), unless DNSG_CS_PAIRED=0 (the scatter form; same state)
        const char *cp = getenv("DNSG_CS_PAIRED");
        h->cs_pair = !(cp && cp[0] == '0');
        const char *e = getenv("DNSG_BAND2");
        const bool pick = (p->precision == DNSG_PREC_FP64 || p->precision == DNSG_PREC_FP32) &&
                          p->force_mode == DNSG_FORCE_CONST;
        h->band2 = h->d.GR == 2 && ((e && e[0]) ? e[0] == '1' : pick);
    }
    const dnsg_dims &d = h->d;
    h->coarse.nx = d.Xc; h->coarse.ny = d.Yc; h->coarse.nz = d.zdc;
    h->lower.nx = d.Xf; h->lower.ny = d.Yf; h->lower.nz = d.zdf - 1;
    h->upper = h->lower;
    dnsg::Grid *gs[3] = {&h->coarse, &h->lower, &h->upper};
    for (auto *g : gs)
        for (int b = 0; b < 2; b++) {
            const size_t bytes = g->qs() * 19 * h->elem;
            HIPCHK(h, hipMalloc(&g->buf[b], bytes));
            HIPCHK(h, hipMemsetAsync(g->buf[b], 0, bytes, h->stream));
        }
    const size_t np = (size_t)d.Xc * d.Yc * NVP;
    for (int l = 0; l < 4; l++) {
        HIPCHK(h, hipMalloc(&h->lev_lo[l], np * 8));
        HIPCHK(h, hipMalloc(&h->lev_up[l], np * 8));
        HIPCHK(h, hipMemsetAsync(h->lev_lo[l], 0, np * 8, h->stream));
        HIPCHK(h, hipMemsetAsync(h->lev_up[l], 0, np * 8, h->stream));
    }
    const std::vector<int8_t> w = wall_codes(p, &d);
    HIPCHK(h, hipMalloc(&h->wall_lo, w.size()));
    HIPCHK(h, hipMalloc(&h->wall_up, w.size()));
    HIPCHK(h, hipMemcpy(h->wall_lo, w.data(), w.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->wall_up, w.data(), w.size(), hipMemcpyHostToDevice));
    h->part_cap = partial_doubles(h);
    HIPCHK(h, hipMalloc(&h->part, h->part_cap * 8));
    {
        const size_t nred =
            ((size_t)d.Yc * d.zdc * cs_chunks(d.Xc) + 2 * (size_t)d.Yf * (d.zdf - 1) * band_cpr_max(h)) * RED_W;
        HIPCHK(h, hipMalloc(&h->red, nred * 8));
        HIPCHK(h, hipMemsetAsync(h->red, 0, nred * 8, h->stream));
        HIPCHK(h, hipMalloc(&h->seg, enc_seg_doubles(h) * 8));
        HIPCHK(h, hipMalloc(&h->dsum, dsum_doubles(h) * 8));
        HIPCHK(h, hipMalloc(&h->fsum, 2 * (size_t)ftcd_blocks(h) * 8));
        HIPCHK(h, hipMalloc(&h->enc_next, 8 * 8));
    }
    HIPCHK(h, hipMalloc(&h->d_force, 4 * 8));
    HIPCHK(h, hipMemsetAsync(h->d_force, 0, 4 * 8, h->stream));
    HIPCHK(h, hipMalloc(&h->d_halt, sizeof(int)));
    HIPCHK(h, hipMemsetAsync(h->d_halt, 0, sizeof(int), h->stream));
    h->halt_flag = h->d_halt;
    if ((rc = noalloc_scratch(h, nullptr))) return rc;
    h->err.reserve(256);  // error texts of dnsg_step_noalloc fit without a reallocation
    const size_t nred = 4 * ((size_t)d.zdc + 2 * (d.zdf - 1)) + 2 * 9 * (size_t)d.NZ + 64;
    HIPCHK(h, hipMalloc(&h->d_red, nred * 8));
    HIPCHK(h, hipMemsetAsync(h->d_red, 0, nred * 8, h->stream));
    HIPCHK(h, hipMalloc(&h->d_acc, 14 * (size_t)d.NZ * 8));
    HIPCHK(h, hipMemsetAsync(h->d_acc, 0, 14 * (size_t)d.NZ * 8, h->stream));
    if ((rc = ensure_diag(h, 64))) return rc;
    if ((rc = ensure_pinned(h, 64 + 9 * (size_t)d.NZ))) return rc;
    h->host_wall_lo = (double *)calloc((size_t)d.Xf * d.Yf * 19, 8);
    h->host_wall_up = (double *)calloc((size_t)d.Xf * d.Yf * 19, 8);
    if (!h->host_wall_lo || !h->host_wall_up) {
        h->err = "host allocation failed";
        return DNSG_ERR_OOM;
    }
    if (p->fltr) {
        HIPCHK(h, hipMalloc(&h->fneq, fltr_fneq_doubles(h) * 8));
        HIPCHK(h, hipMalloc(&h->fwall, fltr_wall_doubles(h) * 8));
        HIPCHK(h, hipMemsetAsync(h->fwall, 0, fltr_wall_doubles(h) * 8, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNSG_OK;
}

extern "C" {

// the filtered injection reads the band wall layers on the device
static int sync_fwall(dnsg_t *h) {
    if (!h->fwall) return DNSG_OK;
    const size_t nw = (size_t)h->d.Xf * h->d.Yf * 19;
    HIPCHK(h, hipMemcpyAsync(h->fwall, h->host_wall_lo, nw * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->fwall + nw, h->host_wall_up, nw * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNSG_OK;
}

}  // extern "C"

// once allocated, fwall is refreshed by every call that sets the state
int dnsg::ensure_fwall(dnsg_t *h) {
    if (h->fwall) return DNSG_OK;
    HIPCHK(h, hipMalloc(&h->fwall, fltr_wall_doubles(h) * 8));
    return sync_fwall(h);
}

extern "C" {

static int set_force(dnsg_t *h, double g) {
    h->force = g;
    HIPCHK(h, hipMemcpyAsync(h->d_force, &h->force, 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DNSG_OK;
}

int dnsg_init_state(dnsg_t *h) {
    if (!h || !h->stream) return DNSG_ERR_ARG;
    h->flux_fused = false;
    HIPCHK(h, launch_init_lattice(h->stream, h, 0, h->coarse.c()));
    HIPCHK(h, launch_init_lattice(h->stream, h, 1, h->lower.c()));
    HIPCHK(h, launch_init_lattice(h->stream, h, 2, h->upper.c()));
    const size_t np = (size_t)h->d.Xc * h->d.Yc * NVP;
    const size_t nw = (size_t)h->d.Xf * h->d.Yf;
    if (h->p.init_kind == DNSG_INIT_REST) {
        // '-s' mode: walls and interface levels stay zero (calloc in main:473-883)
        memset(h->host_wall_lo, 0, nw * 19 * 8);
        memset(h->host_wall_up, 0, nw * 19 * 8);
        h->wall_mass = 0.;
        for (int l = 0; l < 3; l++) {
            HIPCHK(h, hipMemsetAsync(h->lev_lo[l], 0, np * 8, h->stream));
            HIPCHK(h, hipMemsetAsync(h->lev_up[l], 0, np * 8, h->stream));
        }
    } else {
        // wall nodes at rest; three equal levels sampled from the initial coarse state
        static const double wq[19] = {1. / 3., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                      1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                      1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
        double m = 0.;
        for (size_t n = 0; n < nw; n++) {
            double r = 0.;
            for (int a = 0; a < 19; a++) {
                h->host_wall_lo[n * 19 + a] = wq[a];
                h->host_wall_up[n * 19 + a] = wq[a];
                r += wq[a];
            }
            m += r;
        }
        h->wall_mass = 2. * m;
        HIPCHK(h, launch_sample(h->stream, h, h->coarse.c(), h->lev_lo[0], h->lev_up[0]));
        for (int l = 1; l < 3; l++) {
            HIPCHK(h, hipMemcpyAsync(h->lev_lo[l], h->lev_lo[0], np * 8, hipMemcpyDeviceToDevice, h->stream));
            HIPCHK(h, hipMemcpyAsync(h->lev_up[l], h->lev_up[0], np * 8, hipMemcpyDeviceToDevice, h->stream));
        }
    }
    h->ss = 0;
    h->have_state = true;
    int rc = sync_fwall(h);
    if (rc) return rc;
    return set_force(h, h->p.force_mode == DNSG_FORCE_CONST ? h->p.g_const : 0.);
}

// the state of the '-p' read path, in either host layout (lin/uin: newest first)
static int upload_any(dnsg_t *h, const double *coarse, const double *lower, const double *upper,
                      const double *const *lin, const double *const *uin, double force, HostLayout L) {
    if (h) h->flux_fused = false;
    if (!h || !h->stream) return DNSG_ERR_ARG;
    int rc;
    const int zdf = h->d.zdf;
    if (coarse && (rc = put_lattice(h, coarse, h->coarse, h->d.zdc, 0, L))) return rc;
    double wl = 0., wu = 0.;
    bool wall_changed = false;
    if (lower) {
        if ((rc = put_lattice(h, lower, h->lower, zdf, 1, L))) return rc;  // stored m = k-1
        wl = keep_wall(h, lower, 0, h->host_wall_lo, L);
        wall_changed = true;
    } else {
        for (size_t n = 0; n < (size_t)h->d.Xf * h->d.Yf * 19; n++) wl += h->host_wall_lo[n];
    }
    if (upper) {
        if ((rc = put_lattice(h, upper, h->upper, zdf, 0, L))) return rc;  // stored m = k
        wu = keep_wall(h, upper, zdf - 1, h->host_wall_up, L);
        wall_changed = true;
    } else {
        for (size_t n = 0; n < (size_t)h->d.Xf * h->d.Yf * 19; n++) wu += h->host_wall_up[n];
    }
    if (wall_changed) h->wall_mass = wl + wu;
    if (lin && (rc = put_levels(h, lin, h->lev_lo, L))) return rc;
    if (uin && (rc = put_levels(h, uin, h->lev_up, L))) return rc;
    h->ss = 0;
    h->have_state = true;
    if ((rc = sync_fwall(h))) return rc;
    return set_force(h, h->p.force_mode == DNSG_FORCE_CONST ? h->p.g_const : force);
}

int dnsg_upload_state(dnsg_t *h, const double *coarse, const double *lower, const double *upper, const double *lin3,
                      const double *uin3, double force) {
    if (!h) return DNSG_ERR_ARG;
    const size_t np = (size_t)h->d.Xc * h->d.Yc * NVP;
    const double *lin[3] = {lin3, lin3 + np, lin3 + 2 * np}, *uin[3] = {uin3, uin3 + np, uin3 + 2 * np};
    return upload_any(h, coarse, lower, upper, lin3 ? lin : nullptr, uin3 ? uin : nullptr, force, HostLayout{});
}

int dnsg_upload_state_ghosted(dnsg_t *h, const double *ftemp, const double *slftemp, const double *suftemp,
                              const double *const *lin, const double *const *uin, double force) {
    if (lin && (!lin[0] || !lin[1] || !lin[2])) return DNSG_ERR_ARG;
    if (uin && (!uin[0] || !uin[1] || !uin[2])) return DNSG_ERR_ARG;
    HostLayout L;
    L.ghost = true;
    return upload_any(h, ftemp, slftemp, suftemp, lin, uin, force, L);
}

int dnsg_init_uniform(dnsg_t *h, const double *uniform, int xdu, int ydu, int zdu) {
    if (!h || !h->stream || !uniform) return DNSG_ERR_ARG;
    int rc;
    if ((rc = uniform_check(h, xdu, ydu, zdu, &h->err))) return rc;
    h->flux_fused = false;
    const size_t nu = (size_t)xdu * ydu * (zdu + 2) * 19;
    double *du = nullptr;
    HIPCHK(h, hipMalloc(&du, nu * 8));
    hipError_t e = hipMemcpy(du, uniform, nu * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_uniform(h->stream, h, du, xdu, ydu, zdu);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(du);
    HIPCHK(h, e);
    // main:986-987 call init() before solid_init(), so the band wall layers start at
    // rest like every other node (init.c:17-52) and the interpolation leaves them so
    static const double wq[19] = {1. / 3., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
    const size_t nw = (size_t)h->d.Xf * h->d.Yf;
    double m = 0.;
    for (size_t n = 0; n < nw; n++) {
        double r = 0.;
        for (int a = 0; a < 19; a++) {
            h->host_wall_lo[n * 19 + a] = wq[a];
            h->host_wall_up[n * 19 + a] = wq[a];
            r += wq[a];
        }
        m += r;
    }
    h->wall_mass = 2. * m;
    if ((rc = sync_fwall(h))) return rc;
    // inter_init.c:150-257: the k = 1 plane's f_neq and the k = zdc-2 plane's f_neq
    // fill the newest lower / upper levels, but rho and u of both planes are written
    // into the lower one, so it ends with the upper plane's and the upper level keeps
    // zero macros; the older levels are copied over the ghosted (Xc+6) x (Yc+6) planes
    // with a stride of 19 where they have 23, which copies the first 19/23 of each
    // plane's memory and leaves the rest zero
    HIPCHK(h, launch_sample(h->stream, h, h->coarse.c(), h->lev_lo[0], h->lev_up[0]));
    const int Xc = h->d.Xc, Yc = h->d.Yc;
    const size_t np = (size_t)Xc * Yc * NVP;
    std::vector<double> lo(3 * np), up(3 * np);
    double *lop[3] = {lo.data(), lo.data() + np, lo.data() + 2 * np};
    double *upp[3] = {up.data(), up.data() + np, up.data() + 2 * np};
    if ((rc = get_levels(h, lop, h->lev_lo)) || (rc = get_levels(h, upp, h->lev_up))) return rc;
    const size_t lim = (size_t)19 * (Xc + 6) * (Yc + 6);
    for (int i = 0; i < Xc; i++)
        for (int j = 0; j < Yc; j++)
            for (int v = 0; v < NVP; v++) {
                const size_t at = ((size_t)i * Yc + j) * NVP + v;
                if (v < 4) {
                    lo[at] = up[at];
                    up[at] = 0.;
                }
                const bool copied = ((size_t)(i + 2) * (Yc + 6) + (size_t)(j + 2)) * NVP + v < lim;
                lo[np + at] = lo[2 * np + at] = copied ? lo[at] : 0.;
                up[np + at] = up[2 * np + at] = copied ? up[at] : 0.;
            }
    if ((rc = put_levels(h, lop, h->lev_lo)) || (rc = put_levels(h, upp, h->lev_up))) return rc;
    h->ss = 0;
    h->have_state = true;
    return set_force(h, h->p.force_mode == DNSG_FORCE_CONST ? h->p.g_const : 0.);
}

static int download_any(dnsg_t *h, double *coarse, double *lower, double *upper, double *const *lin,
                        double *const *uin, double *force, HostLayout L) {
    if (!h || !h->stream) return DNSG_ERR_ARG;
    int rc;
    const int zdf = h->d.zdf;
    if (coarse && (rc = get_lattice(h, coarse, h->coarse, h->d.zdc, 0, L))) return rc;
    if (lower) {
        if ((rc = get_lattice(h, lower, h->lower, zdf, 1, L))) return rc;
        place_wall(h, lower, 0, h->host_wall_lo, L);
    }
    if (upper) {
        if ((rc = get_lattice(h, upper, h->upper, zdf, 0, L))) return rc;
        place_wall(h, upper, zdf - 1, h->host_wall_up, L);
    }
    if (lin && (rc = get_levels(h, lin, h->lev_lo, L))) return rc;
    if (uin && (rc = get_levels(h, uin, h->lev_up, L))) return rc;
    if (force) {
        HIPCHK(h, hipMemcpyAsync(h->h_pinned, h->d_force, 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        *force = h->h_pinned[0];
    }
    return DNSG_OK;
}

int dnsg_download_state(dnsg_t *h, double *coarse, double *lower, double *upper, double *lin3, double *uin3,
                        double *force) {
    if (!h) return DNSG_ERR_ARG;
    const size_t np = (size_t)h->d.Xc * h->d.Yc * NVP;
    double *lin[3] = {lin3, lin3 + np, lin3 + 2 * np}, *uin[3] = {uin3, uin3 + np, uin3 + 2 * np};
    return download_any(h, coarse, lower, upper, lin3 ? lin : nullptr, uin3 ? uin : nullptr, force, HostLayout{});
}

int dnsg_download_state_ghosted(dnsg_t *h, double *ftemp, double *slftemp, double *suftemp, double *const *lin,
                                double *const *uin, double *force) {
    if (lin && (!lin[0] || !lin[1] || !lin[2])) return DNSG_ERR_ARG;
    if (uin && (!uin[0] || !uin[1] || !uin[2])) return DNSG_ERR_ARG;
    HostLayout L;
    L.ghost = true;
    return download_any(h, ftemp, slftemp, suftemp, lin, uin, force, L);
}

int dnsg_step(dnsg_t *h, int n, dnsg_diag *diag) {
    if (!h || !h->stream || n < 0) return DNSG_ERR_ARG;
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    const int chunk = 256;
    for (int s0 = 0; s0 < n; s0 += chunk) {
        const int m = (n - s0 < chunk) ? n - s0 : chunk;
        if (int rc = run_steps(h, m, diag ? diag + s0 : nullptr, nullptr, nullptr)) return rc;
    }
    return DNSG_OK;
}

int dnsg_step_noalloc(dnsg_t *h, int n, dnsg_diag *diag) {
    if (!h || !h->stream || n < 0) return DNSG_ERR_ARG;
    if (n > DNSG_NOALLOC_MAX_STEPS) {
        h->err = "dnsg_step_noalloc: n exceeds DNSG_NOALLOC_MAX_STEPS (the records of the queued steps "
                 "live in a fixed device array)";
        return DNSG_ERR_ARG;
    }
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    if (n == 0) return DNSG_OK;
    double *base = nullptr;
    int rc;
    if ((rc = noalloc_scratch(h, &base))) return rc;
    const StepBufs B = {base + NOALLOC_REC_OFS, reinterpret_cast<int *>(base), false};
    return run_steps(h, n, diag, nullptr, nullptr, &B);
}

int dnsg_run_timed(dnsg_t *h, int n, dnsg_diag *diag, double *elapsed_ms) {
    return dnsg_run_timed_sections(h, n, diag, elapsed_ms, nullptr, -1);
}

int dnsg_run_timed_sections(dnsg_t *h, int n, dnsg_diag *diag, double *elapsed_ms, double *sec_ms, int group) {
    if (!h || !h->stream || n < 1 || group < -1 || group >= SEC_N) return DNSG_ERR_ARG;
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    SecTimer T(group);  // section events of the timed steps, when sec_ms is given
    float ms = 0.f;
    int rc = run_steps(h, n, diag, &ms, sec_ms ? &T : nullptr);
    if (rc) return rc;
    if (sec_ms && (rc = sec_average(h, T, sec_ms))) return rc;
    if (elapsed_ms) *elapsed_ms = ms;
    return DNSG_OK;
}

int dnsg_band_launch_substeps(const dnsg_t *h, int *substeps) {
    if (!h || !substeps) return DNSG_ERR_ARG;
    *substeps = h->band2 ? 2 : 1;
    return DNSG_OK;
}

int dnsg_kernel_times(dnsg_t *h, int n, double *ms) {
    if (!h || !h->stream || n < 1 || !ms) return DNSG_ERR_ARG;
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    int rc;
    if ((rc = ensure_diag(h, 1))) return rc;
    SecTimer T;
    for (int s = 0; s < n && !rc; s++) rc = do_step(h, h->d_diag, &T);
    if (!rc && !T.ev.empty()) HIPCHK(h, hipEventSynchronize(T.ev.back()));
    if (!rc) rc = sec_average(h, T, ms);
    return rc;
}

// ---- reference-granular entry points ------------------------------------------------

// copy the one diag record of a granular call back; DNSG_OK or a HIP error
static int scalar_call(dnsg_t *h, long ss, dnsg_diag *d) {
    HIPCHK(h, hipMemcpyAsync(h->h_pinned, h->d_diag, sizeof(double) * DIAG_WORDS, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (d) {
        dnsg_diag t;
        fill_diag(h, h->h_pinned, ss, &t);
        d->time = t.time;
    }
    return DNSG_OK;
}

int dnsg_encalc2(dnsg_t *h, long ss, dnsg_diag *d) {
    if (!h || !h->stream) return DNSG_ERR_ARG;
    if (h->flux_fused)  // the previous dnsg_step already formed this state's controller output
        HIPCHK(h, hipMemcpyAsync(h->d_diag + 1, h->enc_next, 4 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    else
        HIPCHK(h, launch_encalc2(h->stream, h, h->d_diag));
    if (int rc = scalar_call(h, ss, d)) return rc;
    const double *rec = h->h_pinned;
    h->force = rec[1];
    if (d) { d->force = rec[1]; d->momentum = rec[2]; d->density = rec[3]; d->flowrate = rec[4]; }
    return DNSG_OK;
}

int dnsg_wshearstr(dnsg_t *h, long ss, dnsg_diag *d) {
    if (!h || !h->stream) return DNSG_ERR_ARG;
    HIPCHK(h, launch_wshear(h->stream, h, h->d_diag));
    if (int rc = scalar_call(h, ss, d)) return rc;
    const double *rec = h->h_pinned;
    if (d) { d->str_up = rec[5]; d->str_down = rec[6]; d->stress = rec[7]; d->pgrad = rec[8]; }
    if (isnan(rec[7])) {
        h->err = "NaN in wall shear stress";
        return DNSG_ERR_NAN;
    }
    return DNSG_OK;
}

// The granular calls change the state outside dnsg_step: the next step's flux
// controller then sums the lattices again (flux_fused = false).
int dnsg_xcomputations(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    h->flux_fused = false;
    return do_xcomputations(h);
}

int dnsg_xfcomputations(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    h->flux_fused = false;
    return do_xfcomputations(h);
}

int dnsg_ctfdtransfer(dnsg_t *h, int iter) {
    if (!h || iter < 1 || iter > h->d.GR) return DNSG_ERR_ARG;
    h->flux_fused = false;
    HIPCHK(h, launch_interface(h->stream, h, iter));
    return DNSG_OK;
}

int dnsg_ftcdtransfer(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    h->flux_fused = false;
    HIPCHK(h, launch_ftcd(h->stream, h));
    return DNSG_OK;
}

int dnsg_exchange(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    h->flux_fused = false;
    h->coarse.swap();
    h->ss++;
    return DNSG_OK;
}

int dnsg_fexchange(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    h->flux_fused = false;
    h->lower.swap();
    h->upper.swap();
    return DNSG_OK;
}

// ---- reductions and snapshots ----------------------------------------------------------

int dnsg_mass(dnsg_t *h, double *avg) {
    if (!h || !avg) return DNSG_ERR_ARG;
    HIPCHK(h, launch_mass(h->stream, h, h->d_red));
    HIPCHK(h, hipMemcpyAsync(h->h_pinned, h->d_red, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *avg = h->h_pinned[0];
    return DNSG_OK;
}

static double *stats_scratch(dnsg_t *h) { return h->d_red + 4 * ((size_t)h->d.zdc + 2 * (h->d.zdf - 1)); }

int dnsg_statistics(dnsg_t *h, double *out) {
    if (!h || !out) return DNSG_ERR_ARG;
    double *dst = stats_scratch(h);
    const size_t n = 9 * (size_t)h->d.NZ;
    HIPCHK(h, hipMemsetAsync(dst, 0, n * 8, h->stream));
    HIPCHK(h, launch_statistics(h->stream, h, dst));
    HIPCHK(h, hipMemcpyAsync(h->h_pinned, dst, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    memcpy(out, h->h_pinned, n * 8);
    return DNSG_OK;
}

int dnsg_stats_accumulate(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    double *dst = stats_scratch(h);
    const int NZ = h->d.NZ;
    HIPCHK(h, hipMemsetAsync(dst, 0, 9 * (size_t)NZ * 8, h->stream));
    HIPCHK(h, launch_statistics(h->stream, h, dst));
    HIPCHK(h, launch_fold_stats(h->stream, h->d_acc, dst, NZ));
    h->acc_count++;
    return DNSG_OK;
}

int dnsg_stats_average(dnsg_t *h, double *out, long *count) {
    if (!h || !out) return DNSG_ERR_ARG;
    const int NZ = h->d.NZ;
    const size_t n = 14 * (size_t)NZ;
    if (int rc = ensure_pinned(h, n)) return rc;
    HIPCHK(h, hipMemcpyAsync(h->h_pinned, h->d_acc, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const double T = h->acc_count > 0 ? (double)h->acc_count : 1.;
    const double *a = h->h_pinned;
    for (int k = 0; k < NZ; k++) {
        auto m = [&](int r) { return a[(size_t)r * NZ + k] / T; };
        const double U = m(0), V = m(1), W = m(2), D = m(7);
        auto root = [](double v) { return v > 0. ? sqrt(v) : 0.; };
        out[0 * NZ + k] = U;
        out[1 * NZ + k] = V;
        out[2 * NZ + k] = W;
        out[3 * NZ + k] = root(m(3) + (m(9) - U * U));
        out[4 * NZ + k] = root(m(4) + (m(10) - V * V));
        out[5 * NZ + k] = root(m(5) + (m(11) - W * W));
        out[6 * NZ + k] = m(6) + (m(12) - U * W);
        out[7 * NZ + k] = 1. + D;
        out[8 * NZ + k] = root(m(8) + (m(13) - D * D));
    }
    if (count) *count = h->acc_count;
    return DNSG_OK;
}

int dnsg_stats_reset(dnsg_t *h) {
    if (!h) return DNSG_ERR_ARG;
    HIPCHK(h, hipMemsetAsync(h->d_acc, 0, 14 * (size_t)h->d.NZ * 8, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->acc_count = 0;
    return DNSG_OK;
}

int dnsg_velocity_snapshot(dnsg_t *h, int grid, double *vel, double *den) {
    if (!h || !h->stream || grid < 0 || grid > 2) return DNSG_ERR_ARG;
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    // vel_print.c:86-117 on the device, then one copy of 4 values per node
    int rc = ensure_fwall(h);
    if (rc) return rc;
    const int X = grid == 0 ? h->d.Xc : h->d.Xf, Y = grid == 0 ? h->d.Yc : h->d.Yf;
    const int Z = grid == 0 ? h->d.zdc : h->d.zdf;
    const size_t nn = (size_t)X * Y * Z;
    double *dv = nullptr;
    HIPCHK(h, hipMalloc(&dv, nn * 4 * 8));
    hipError_t e = launch_snapshot(h, grid, 0, dv, dv + 3 * nn);
    if (e == hipSuccess && vel) e = hipMemcpyAsync(vel, dv, nn * 3 * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && den) e = hipMemcpyAsync(den, dv + 3 * nn, nn * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    else (void)hipStreamSynchronize(h->stream);
    (void)hipFree(dv);
    HIPCHK(h, e);
    return DNSG_OK;
}

}  // extern "C"
