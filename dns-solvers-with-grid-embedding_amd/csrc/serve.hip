// serve.hip -- the serving queue of steady slip-length solves (dnsg_queue_*).
//
// A request is a whole grid-embedded channel (its own sizes, gratio, wall
// pattern, relaxation time and body force).  Admitted requests become resident
// solver handles on the queue's stream; each coarse step of the resident set is
// one batched launch per kernel and precision (tables of up to BATCH_MAX items
// per launch: lattices, bands or solves, by kernel), so small channels share the
// GPU instead of each leaving most of it idle.  The per-solve arithmetic is the
// single-solve step's, unchanged: a solve gives the same bits alone or in any mix.
//
// Convergence: every check_every steps the bulk velocity U = <rho u_x>/<rho>
// (encalc2.c:109-131 weights) of every resident solve is copied back in one
// transfer; a solve retires when |U(t) - U(t - window)| <= tol |U(t)|.  FP32
// storage leaves U wandering at about 1e-5 relative once the flow is steady, so
// a shifted-tier solve also retires (status 2) when U has stayed within
// floor_tol |U| of its values 1..5 windows back and has moved both ways.
// Slip length: with walls halfway between nodes the channel height is
// H = (NZs - 2)/GR coarse cells; a no-slip channel under body force G carries
// U0 = G H^2 / (12 nu_c); equal slip b on both walls raises it to
// U0 (1 + 6 b / H), so b = (H/6)(U/U0 - 1).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace dnsg;

namespace {

struct Solve {
    uint64_t id = 0;
    uint64_t serial = 0;  // admission number, unique over the queue's life
    dnsg_params p;
    dnsg_dims d;
    size_t bytes = 0;
    double t_arrival = 0., t_admit = 0.;
    dnsg_ctx *h = nullptr;
    std::vector<double> hist;  // U at each check since admission
    int slot = -1;
};

// Device bytes of one resident solve, estimated before it is created: what create_ctx
// allocates, each buffer of 1 MiB or more rounded up to the runtime's 2 MiB allocation
// granularity, plus 2 MiB for the small buffers' pool.  The queue replaces it by the
// measured footprint when that is larger (admit).
constexpr size_t GRAN = (size_t)2 << 20;
inline size_t gran(size_t b) { return b >= ((size_t)1 << 20) ? (b + GRAN - 1) / GRAN * GRAN : b; }
size_t solve_bytes(const dnsg_params *p, const dnsg_dims &d) {
    const size_t elem = p->precision == DNSG_PREC_FP64 ? 8 : (p->precision == DNSG_PREC_FP32 ? 4 : 2);
    const size_t nc = (size_t)d.Xc * d.Yc * d.zdc, nb = (size_t)d.Xf * d.Yf * (d.zdf - 1);
    size_t b = 2 * gran(nc * 19 * elem) + 4 * gran(nb * 19 * elem);  // A/B lattices
    b += 8 * gran((size_t)d.Xc * d.Yc * 23 * 8);            // interface levels (3 + a spare per side)
    b += 2 * gran((size_t)d.Xf * d.Yf);                     // wall codes
    const size_t rows = (size_t)d.Yc * d.zdc + 2 * (size_t)d.Yf * (d.zdf - 1);
    b += rows * 5 * 8 + (size_t)d.Yf * 4 * 8;               // explicit-pass partials
    b += ((size_t)d.Yc * d.zdc * cs_chunks(d.Xc) + 2 * (size_t)d.Yf * (d.zdf - 1) * (size_t)((d.Xf + B2_TX - 1) / B2_TX)) * RED_W * 8 * 2;
    b += (4 * ((size_t)d.zdc + 2 * (d.zdf - 1)) + 2 * 9 * (size_t)d.NZ + 64) * 8 + 14 * (size_t)d.NZ * 8;
    b += 64 * 9 * 8 + 4 * 8;
    return b + GRAN;  // the small buffers' pool
}

}  // namespace

struct dnsg_queue {
    dnsg_queue_opts o;
    std::string err;
    hipStream_t stream = nullptr;
    std::deque<Solve *> pending;
    std::vector<Solve *> active;
    std::deque<dnsg_slip_result> finished;
    size_t budget = 0, resident = 0;
    double *d_u = nullptr, *h_u = nullptr;  // 2 doubles per slot: <rho u_x>, <rho>
    std::chrono::steady_clock::time_point t0;
    uint64_t admitted = 0;
    // One check interval of the resident set as a HIP graph: captured once the set
    // has stayed the same for a whole interval, replayed while it stays so.  Every
    // pointer the launches take returns to its value after an interval (the buffer
    // swaps and the four-level rotation have periods dividing 4, and intervals are
    // multiples of 4 steps), so one capture serves every later interval of that set.
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint64_t> gset;     // serials of the set the graph belongs to
    std::vector<uint64_t> last_set;  // serials of the set of the previous interval
    long graph_launches = 0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

#define QCHK(q, expr)                                                                \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            (q)->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return DNSG_ERR_HIP;                                                     \
        }                                                                            \
    } while (0)

namespace {

dnsg_slip_result make_result(const dnsg_queue *q, const Solve *s, int status) {
    dnsg_slip_result r;
    memset(&r, 0, sizeof r);
    r.id = s->id;
    r.status = status;
    r.precision = s->p.precision;
    r.steps = (long)s->hist.size() * q->o.check_every;
    r.ubulk = s->hist.empty() ? 0. : s->hist.back();
    const double H = (s->d.NZs - 2.) / s->d.GR;
    const double nu = (s->d.tau_c - 0.5) / 3.;
    r.u_noslip = s->p.g_const * H * H / (12. * nu);
    r.slip_length = (H / 6.) * (r.ubulk / r.u_noslip - 1.);
    r.t_arrival = s->t_arrival;
    r.t_admit = s->t_admit;
    r.t_done = q->now();
    return r;
}

void retire(dnsg_queue *q, size_t i, int status) {
    Solve *s = q->active[i];
    q->finished.push_back(make_result(q, s, status));
    dnsg_destroy(s->h);
    q->resident -= s->bytes;
    delete s;
    q->active.erase(q->active.begin() + (long)i);
}

// first fit in arrival order
int admit(dnsg_queue *q) {
    for (auto it = q->pending.begin(); it != q->pending.end();) {
        if ((int)q->active.size() >= q->o.max_active) break;
        Solve *s = *it;
        if (q->resident + s->bytes > q->budget) {
            ++it;
            continue;
        }
        // the device bytes the solve actually takes (the runtime rounds and pools
        // allocations), measured around its creation; accounted as the larger of that
        // and the estimate the admission used
        size_t fr0 = 0, fr1 = 0, tot = 0;
        const bool measured = hipMemGetInfo(&fr0, &tot) == hipSuccess;
        int rc = create_ctx(&s->p, q->stream, &s->h);
        if (!rc) rc = dnsg_init_state(s->h);
        if (!rc && measured && hipMemGetInfo(&fr1, &tot) == hipSuccess && fr0 > fr1)
            s->bytes = std::max(s->bytes, fr0 - fr1);
        if (rc) {
            q->err = std::string("admitting solve: ") + (s->h ? dnsg_last_error(s->h) : "");
            if (s->h) dnsg_destroy(s->h);
            s->h = nullptr;
            dnsg_slip_result r = make_result(q, s, rc);
            q->finished.push_back(r);
            delete s;
            it = q->pending.erase(it);
            continue;
        }
        s->t_admit = q->now();
        s->serial = ++q->admitted;
        q->resident += s->bytes;
        q->active.push_back(s);
        it = q->pending.erase(it);
    }
    for (size_t i = 0; i < q->active.size(); i++) q->active[i]->slot = (int)i;
    return DNSG_OK;
}

// one coarse step of every resident solve (the kernel sequence of the single-
// solve step in constant-force mode, without the wall-stress sums); the bulk
// sums are taken only when want_flux (the step before a check)
int tick(dnsg_queue *q, bool want_flux) {
    for (int prec = 0; prec < 3; prec++) {
        std::vector<dnsg_ctx *> g;
        std::vector<int> slots;
        int maxGR = 0;
        for (Solve *s : q->active)
            if (s->p.precision == prec) {
                g.push_back(s->h);
                slots.push_back(s->slot);
                maxGR = std::max(maxGR, s->d.GR);
            }
        if (g.empty()) continue;
        const int n = (int)g.size();
        std::vector<LatArgs> L(3 * (size_t)n);
        for (int iter = 1; iter <= maxGR; iter++) {
            std::vector<dnsg_ctx *> sub;
            for (dnsg_ctx *h : g)
                if (h->d.GR >= iter) sub.push_back(h);
            const int m = (int)sub.size();
            int nl = 0;
            if (iter == 1) {  // coarse grids and first band substep in one launch set
                for (int i = 0; i < n; i++) {
                    L[nl] = coarse_args(g[i]);
                    if (want_flux) L[nl].red = red_coarse(g[i]);
                    nl++;
                }
            }
            for (int i = 0; i < m; i++) {
                band_args(sub[i], &L[nl]);
                if (iter == 1 && want_flux) {
                    L[nl].red = red_lower(sub[i]);
                    L[nl + 1].red = red_upper(sub[i]);
                }
                nl += 2;
            }
            QCHK(q, launch_collide_stream(q->stream, L.data(), nl, prec));
            if (iter == 1) {
                for (dnsg_ctx *h : g) rotate_levels(h);
                QCHK(q, launch_sample_many(q->stream, g.data(), n));
                if (want_flux) QCHK(q, launch_flux_many(q->stream, g.data(), n, q->d_u, slots.data()));
            }
            for (int GR = 1; GR <= 4; GR++) {
                std::vector<dnsg_ctx *> same;
                for (dnsg_ctx *h : sub)
                    if (h->d.GR == GR) same.push_back(h);
                if (same.empty()) continue;
                QCHK(q, launch_interface_many(q->stream, same.data(), (int)same.size(), iter));
                if (iter == GR) QCHK(q, launch_ftcd_many(q->stream, same.data(), (int)same.size()));
            }
            for (dnsg_ctx *h : sub) {
                h->lower.swap();
                h->upper.swap();
            }
        }
        for (dnsg_ctx *h : g) {
            h->coarse.swap();
            h->ss++;
        }
    }
    return DNSG_OK;
}

// Steady at the storage precision: U has stayed within bound of its values 1..5
// windows back, and the 5 window-to-window changes do not all have one sign (a
// decaying transient moves one way; rounding noise does not).
bool at_floor(const std::vector<double> &h, size_t W, double bound) {
    const size_t e = h.size() - 1;
    const double U = h[e];
    int up = 0, down = 0;
    for (size_t j = 1; j <= 5; j++) {
        if (fabs(U - h[e - j * W]) > bound) return false;
        const double d = h[e - (j - 1) * W] - h[e - j * W];
        up += d > 0.;
        down += d < 0.;
    }
    return up > 0 && down > 0;
}

// read every resident solve's bulk velocity and retire the steady ones
int check(dnsg_queue *q) {
    const size_t na = q->active.size();
    if (!na) return DNSG_OK;
    QCHK(q, hipMemcpyAsync(q->h_u, q->d_u, 2 * na * sizeof(double), hipMemcpyDeviceToHost, q->stream));
    QCHK(q, hipStreamSynchronize(q->stream));
    const size_t W = (size_t)(q->o.window / q->o.check_every);
    for (size_t i = na; i-- > 0;) {
        Solve *s = q->active[i];
        const double U = q->h_u[2 * s->slot] / q->h_u[2 * s->slot + 1];
        s->hist.push_back(U);
        const long steps = (long)s->hist.size() * q->o.check_every;
        const size_t nh = s->hist.size();
        if (!isfinite(U)) {
            retire(q, i, DNSG_ERR_NAN);
        } else if (nh > W && fabs(U - s->hist[nh - 1 - W]) <= q->o.tol * fabs(U)) {
            retire(q, i, 0);
        } else if (s->p.precision != DNSG_PREC_FP64 && q->o.floor_tol > 0. && nh > 5 * W &&
                   at_floor(s->hist, W, q->o.floor_tol * fabs(U))) {
            retire(q, i, 2);
        } else if (steps >= q->o.max_steps) {
            retire(q, i, 1);
        }
    }
    return DNSG_OK;
}

void drop_graph(dnsg_queue *q) {
    if (q->gexec) (void)hipGraphExecDestroy(q->gexec);
    q->gexec = nullptr;
    q->gset.clear();
}

// one check interval of the resident set: check_every ticks, the last with the flux
// sums; eager launches, or the interval's graph (captured here when the set was
// already resident for the whole previous interval)
int interval(dnsg_queue *q) {
    const int ce = q->o.check_every;
    std::vector<uint64_t> set;
    set.reserve(q->active.size());
    for (const Solve *s : q->active) set.push_back(s->serial);
    const bool stable = set == q->last_set;
    q->last_set = set;
    if (q->gexec && set != q->gset) drop_graph(q);
    if (q->gexec) {
        QCHK(q, hipGraphLaunch(q->gexec, q->stream));
        q->graph_launches++;
        for (Solve *s : q->active) s->h->ss += ce;
        return DNSG_OK;
    }
    if (!(q->use_graph && stable && ce % 4 == 0)) {
        int rc = DNSG_OK;
        for (int k = 0; k < ce && !rc; k++) rc = tick(q, k == ce - 1);
        return rc;
    }
    // capture: the ticks record their launches (and advance the host pointers through
    // a whole period) instead of running them; then the graph runs the interval
    QCHK(q, hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal));
    int rc = DNSG_OK;
    for (int k = 0; k < ce && !rc; k++) rc = tick(q, k == ce - 1);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(q->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) q->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return rc ? rc : DNSG_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&q->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        q->gexec = nullptr;
        q->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
        return DNSG_ERR_HIP;
    }
    q->gset = set;
    QCHK(q, hipGraphLaunch(q->gexec, q->stream));
    q->graph_launches++;
    return DNSG_OK;
}

}  // namespace

extern "C" {

int dnsg_queue_opts_default(dnsg_queue_opts *o) {
    if (!o) return DNSG_ERR_ARG;
    o->hbm_budget = 0;
    o->max_active = 64;
    o->check_every = 100;
    o->window = 1000;
    o->tol = 1e-9;
    o->max_steps = 200000;
    o->floor_tol = 3e-5;
    return DNSG_OK;
}

int dnsg_queue_create(const dnsg_queue_opts *o, dnsg_queue_t **out) {
    if (!o || !out) return DNSG_ERR_ARG;
    *out = nullptr;
    dnsg_queue *q = new (std::nothrow) dnsg_queue();
    if (!q) return DNSG_ERR_OOM;
    *out = q;
    q->o = *o;
    q->t0 = std::chrono::steady_clock::now();
    if (o->max_active < 1 || o->max_active > DNSG_QUEUE_MAX_ACTIVE || o->check_every < 1 || o->window < o->check_every ||
        o->window % o->check_every || !(o->tol >= 0.) || !(o->floor_tol >= 0.) || o->max_steps < o->check_every) {
        q->err = "invalid queue options";
        return DNSG_ERR_ARG;
    }
    {
        // DNSG_QUEUE_GRAPH=0 in the environment runs every interval with eager launches
        const char *e = getenv("DNSG_QUEUE_GRAPH");
        q->use_graph = !(e && e[0] == '0');
    }
    QCHK(q, hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
    QCHK(q, hipMalloc(&q->d_u, 2 * sizeof(double) * DNSG_QUEUE_MAX_ACTIVE));
    QCHK(q, hipHostMalloc((void **)&q->h_u, 2 * sizeof(double) * DNSG_QUEUE_MAX_ACTIVE, hipHostMallocDefault));
    if (o->hbm_budget) {
        q->budget = o->hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        QCHK(q, hipMemGetInfo(&fr, &tot));
        const size_t reserve = (size_t)4 << 30;
        q->budget = fr > reserve ? fr - reserve : 0;
    }
    return DNSG_OK;
}

int dnsg_queue_destroy(dnsg_queue_t *q) {
    if (!q) return DNSG_OK;
    if (q->stream) (void)hipStreamSynchronize(q->stream);
    drop_graph(q);
    for (Solve *s : q->active) {
        dnsg_destroy(s->h);
        delete s;
    }
    for (Solve *s : q->pending) delete s;
    if (q->d_u) (void)hipFree(q->d_u);
    if (q->h_u) (void)hipHostFree(q->h_u);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
    return DNSG_OK;
}

const char *dnsg_queue_last_error(const dnsg_queue_t *q) { return q ? q->err.c_str() : "null queue"; }

double dnsg_queue_clock(const dnsg_queue_t *q) { return q ? q->now() : 0.; }

size_t dnsg_solve_bytes(const dnsg_params *p) {
    dnsg_dims d;
    if (!p || dnsg_dims_from_params(p, &d)) return 0;
    return solve_bytes(p, d);
}

int dnsg_queue_submit(dnsg_queue_t *q, const dnsg_params *p, uint64_t id, double t_arrival) {
    if (!q || !p) return DNSG_ERR_ARG;
    Solve *s = new (std::nothrow) Solve();
    if (!s) return DNSG_ERR_OOM;
    s->id = id;
    s->p = *p;
    int rc = dnsg_dims_from_params(p, &s->d);
    if (rc || p->force_mode != DNSG_FORCE_CONST || !(p->g_const > 0.)) {
        q->err = rc ? "invalid solver parameters" : "queued solves need force_mode = DNSG_FORCE_CONST and g_const > 0";
        delete s;
        return DNSG_ERR_ARG;
    }
    if (p->fltr) {  // the batched ticks run the plain injection only
        q->err = "queued solves take the plain injection (fltr = 0)";
        delete s;
        return DNSG_ERR_ARG;
    }
    s->bytes = solve_bytes(p, s->d);
    if (s->bytes > q->budget) {
        q->err = "solve does not fit the queue's device-memory budget";
        delete s;
        return DNSG_ERR_ARG;
    }
    s->t_arrival = t_arrival < 0. ? q->now() : t_arrival;
    q->pending.push_back(s);
    return DNSG_OK;
}

int dnsg_queue_advance(dnsg_queue_t *q, int steps, dnsg_slip_result *done, int cap, int *ndone) {
    if (!q || steps < 0 || cap < 0 || (cap && !done) || !ndone) return DNSG_ERR_ARG;
    *ndone = 0;
    int rc = admit(q);
    const int ce = q->o.check_every;
    for (int s = 0; !rc && s < steps && !q->active.empty(); s += ce) {
        rc = interval(q);
        if (!rc) rc = check(q);
        if (!rc) rc = admit(q);
    }
    while (*ndone < cap && !q->finished.empty()) {
        done[(*ndone)++] = q->finished.front();
        q->finished.pop_front();
    }
    return rc;
}

int dnsg_queue_graph_launches(const dnsg_queue_t *q, long *launches) {
    if (!q || !launches) return DNSG_ERR_ARG;
    *launches = q->graph_launches;
    return DNSG_OK;
}

int dnsg_queue_info(const dnsg_queue_t *q, int *pending, int *active, size_t *resident_bytes) {
    if (!q) return DNSG_ERR_ARG;
    if (pending) *pending = (int)q->pending.size();
    if (active) *active = (int)q->active.size();
    if (resident_bytes) *resident_bytes = q->resident;
    return DNSG_OK;
}

}  // extern "C"
