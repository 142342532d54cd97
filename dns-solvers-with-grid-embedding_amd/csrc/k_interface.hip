// k_interface.hip -- the grid-embedding interface between the coarse grid and
// the two wall bands.
//
//   sample     coarse k = 1 / zdc-2 after streaming -> (rho, u, f - feq) planes
//              (xcomputations.c:558-625)
//   iface      per substep and band: cubic Hermite in time of the three newest
//              planes (tinterpolate_hrmt.c:16-55), 4th-order d/dx, d/dy,
//              d2/dxdy (derivative.c:17-89), the bicubic patch of each coarse
//              cell (bcucof_opt.c:16-62) evaluated at its GR x GR fine columns,
//              and f = feq + tau_f/(GR tau_c) f_neq on the band interface layer
//              (ctfdtransfer_opt.c:17-175) -- one LDS-tiled kernel
//   ftcd       fine -> coarse injection at coincident nodes,
//              f = feq + GR tau_c/tau_f f_neq (ftcdtransfer.c:16-118)
// Planes are periodic in x and y and laid out [value][Yc][Xc].
// FP64 tier: FP64 planes, every expression in the reference's operation order,
// so the results are bitwise those of the reference.
// Shifted tiers (lattices hold g = f - w): FP32 planes and arithmetic, slot 0
// holds rho - 1 instead of rho, and f_neq = g - (feq - w).  Their bicubic patch is
// one compile-time weight matrix applied with FMAs instead of the reference's
// coefficient-then-Horner order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {


// ---- sampling --------------------------------------------------------------------
struct SampleItem {
    const void *cs;   // coarse post-stream lattice
    void *plo, *pup;  // newest plane level of each side
    int nx, ny, k_up; // coarse extents; layers 1 and k_up = zdc-2 are sampled
};

// item i: workgroups [0, nb) sample the lower side, [nb, 2 nb) the upper side
template <class S, class P>
__global__ __launch_bounds__(256) void k_sample(Batch<SampleItem> B) {
    if (halted(B.halt)) return;
    const int it = batch_find(B.first, B.n, blockIdx.x);
    const SampleItem &A = B.it[it];
    const int n = A.nx * A.ny, nb = (n + 255) / 256;
    const int lb = blockIdx.x - B.first[it];
    const int side = lb >= nb ? 1 : 0;
    const int id = (lb - side * nb) * 256 + threadIdx.x;
    if (id >= n) return;
    const int k = side == 0 ? 1 : A.k_up;
    P *pl = static_cast<P *>(side == 0 ? A.plo : A.pup);
    const S *cs = static_cast<const S *>(A.cs);
    const int nx = A.nx;
    const size_t row = (size_t)k * A.ny + id / nx;
    const int x = id % nx;
    P f[Q];
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = (P)load_s(cs + lat_at(q, row, x, nx));
    });
    if constexpr (sizeof(S) == 8) {
        const Macro64 m = macro64(f);
        pl[0 * n + id] = m.dens;
        pl[1 * n + id] = m.u0;
        pl[2 * n + id] = m.u1;
        pl[3 * n + id] = m.u2;
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            pl[(4 + q) * n + id] = f[q] - feq64<q>(m.dens, m.u0, m.u1, m.u2, m.v);
        });
    } else {
        const Macro32 m = macro32(f);
        pl[0 * n + id] = m.drho;
        pl[1 * n + id] = m.u0;
        pl[2 * n + id] = m.u1;
        pl[3 * n + id] = m.u2;
        unroll<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            pl[(4 + q) * n + id] = f[q] - geq32<q>(m);
        });
    }
}

// fill a batch table from items [i0, i0 + n) and return its workgroup count
template <class A, class F>
static int fill_batch(Batch<A> &B, const A *items, int n, F blocks_of) {
    B.n = n;
    int blocks = 0;
    for (int i = 0; i < BATCH_MAX; i++) {
        B.it[i] = items[i < n ? i : 0];
        B.first[i] = blocks;
        if (i < n) blocks += blocks_of(B.it[i]);
    }
    B.first[BATCH_MAX] = blocks;
    return blocks;
}

static hipError_t sample_items(hipStream_t s, int prec, const SampleItem *items, int n, const int *halt) {
    for (int i0 = 0; i0 < n; i0 += BATCH_MAX) {
        Batch<SampleItem> B;
        B.halt = halt;
        const int blocks = fill_batch(B, items + i0, std::min(BATCH_MAX, n - i0),
                                      [](const SampleItem &a) { return 2 * ((a.nx * a.ny + 255) / 256); });
        if (prec == DNSG_PREC_FP64)
            hipLaunchKernelGGL((k_sample<double, double>), dim3(blocks), dim3(256), 0, s, B);
        else if (prec == DNSG_PREC_FP32)
            hipLaunchKernelGGL((k_sample<float, float>), dim3(blocks), dim3(256), 0, s, B);
        else
            hipLaunchKernelGGL((k_sample<__half, float>), dim3(blocks), dim3(256), 0, s, B);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sample(hipStream_t s, dnsg_ctx *h, const void *coarse_state, void *lo, void *up) {
    const Grid &g = h->coarse;
    const SampleItem a = {coarse_state, lo, up, g.nx, g.ny, g.nz - 2};
    return sample_items(s, h->p.precision, &a, 1, h->halt());
}

hipError_t launch_sample_many(hipStream_t s, dnsg_ctx *const *hs, int n) {
    std::vector<SampleItem> v(n);
    for (int i = 0; i < n; i++) {
        const Grid &g = hs[i]->coarse;
        v[i] = {g.n(), hs[i]->lev_lo[0], hs[i]->lev_up[0], g.nx, g.ny, g.nz - 2};
    }
    return n ? sample_items(s, hs[0]->p.precision, v.data(), n, nullptr) : hipSuccess;
}

// ---- helpers ---------------------------------------------------------------------------
template <class P>
struct HermiteW {
    P h00, h10, h01, h11;
    int exact;  // every weight is 0 or a power of two, so its products are exact (every
                // substep at gratio 1 and 2, substeps 2 and 4 at gratio 4)
};

// a b + c with one rounding, used only where the product a b is exact (a is 0 or a
// power of two): then it equals round(round(a b) + c), the reference's unfused form,
// bit for bit, in one instruction
__device__ __forceinline__ double fma_exact(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_exact(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

constexpr bool pow2_int(int w) { return w > 0 && (w & (w - 1)) == 0; }
// i / g is a power of two (i > 0)
constexpr bool frac_pow2(int i, int g) { return i > 0 && g % i == 0 && pow2_int(g / i); }

template <class P>
__device__ __forceinline__ P d4(P fm2, P fm1, P fp1, P fp2) {
    return (P)(2. / 3.) * (fp1 - fm1) - (P)(1. / 12.) * (fp2 - fm2);
}

// the bicubic interpolation matrix of bcucof_opt.c:18-35 (Press et al., bcucof)
constexpr int WT[16][16] = {
    {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
    {-3, 0, 0, 3, 0, 0, 0, 0, -2, 0, 0, -1, 0, 0, 0, 0},
    {2, 0, 0, -2, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0},
    {0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0},
    {0, 0, 0, 0, -3, 0, 0, 3, 0, 0, 0, 0, -2, 0, 0, -1},
    {0, 0, 0, 0, 2, 0, 0, -2, 0, 0, 0, 0, 1, 0, 0, 1},
    {-3, 3, 0, 0, -2, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, -3, 3, 0, 0, -2, -1, 0, 0},
    {9, -9, 9, -9, 6, 3, -3, -6, 6, -6, -3, 3, 4, 2, 1, 2},
    {-6, 6, -6, 6, -4, -2, 2, 4, -3, 3, 3, -3, -2, -1, -1, -2},
    {2, -2, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 2, -2, 0, 0, 1, 1, 0, 0},
    {-6, 6, -6, 6, -3, -3, 3, 3, -4, 4, 2, -2, -2, -2, -1, -1},
    {4, -4, 4, -4, 2, 2, -2, -2, 2, -2, -2, 2, 1, 1, 1, 1}};

constexpr int wt_first(int i) {
    for (int k = 0; k < 16; k++)
        if (WT[i][k] != 0) return k;
    return 16;
}

// row i of c = WT x, summed in k order with the zero weights dropped.  The
// reference's sum starts from 0 (bcucof_opt.c:52-55); starting from the first
// term instead can change only the sign of an exactly zero partial sum, which no
// nonzero result sees.
template <int i, class P>
__device__ __forceinline__ P wt_row(const P *xv) {
    P xx;
    unroll<0, 16>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int w = WT[i][k], aw = w < 0 ? -w : w;
        if constexpr (w != 0) {
            if constexpr (k == wt_first(i)) xx = (P)w * xv[k];
            else if constexpr (aw != 1 && pow2_int(aw)) xx = fma_exact((P)w, xv[k], xx);
            else xx = xx + (P)w * xv[k];
        }
    });
    return xx;
}

// The bicubic patch of one cell evaluated at its GR x GR fine points (t, u) =
// (ii, jj)/GR: coefficients c = WT x (bcucof_opt.c:47-61) and the Horner form of
// ctfdtransfer_opt.c:59-62,
//   r = c12 + X3 u, then r = (t r + X_nn u) + c_4nn for nn = 2, 1, 0,
//   X_nn = (c_4nn+3 u + c_4nn+2) u + c_4nn+1,
// with that grouping kept.  Each block of four coefficients is formed just before
// the Horner step that takes it, so only one block is live at a time.  Where t or
// u is 0 the products with it are dropped: for finite coefficients they only add a
// signed zero, which leaves every nonzero result bit-identical (an exactly zero
// f_neq or velocity gives the same stored f, f = feq + rescale f_neq with feq > 0).
// weight of x_k in the patch value at fine point pp: sum over the coefficients
// c_(4a+b) = (WT x)_(4a+b) of t^a u^b, formed in double at compile time
constexpr double bicubic_w(int GR, int pp, int k) {
    const double t = double(pp / GR) / GR, u = double(pp % GR) / GR;
    double acc = 0., ta = 1.;
    for (int a = 0; a < 4; a++) {
        double ub = 1.;
        for (int b = 0; b < 4; b++) {
            acc += ta * ub * WT[4 * a + b][k];
            ub *= u;
        }
        ta *= t;
    }
    return (acc < 1e-12 && acc > -1e-12) ? 0. : acc;
}
constexpr int bicubic_first(int GR, int pp) {
    for (int k = 0; k < 16; k++)
        if (bicubic_w(GR, pp, k) != 0.) return k;
    return 16;
}

template <int GR, class P>
__device__ __forceinline__ void bicubic_patch(const P *xv, P *r) {
    if constexpr (sizeof(P) == 4) {
        // shifted tiers (FP32 arithmetic, compared within a tolerance): the patch is
        // linear in x, so each fine value is one FMA chain over the nonzero weights
        unroll<0, GR * GR>([&](auto pc) {
            constexpr int pp = decltype(pc)::value, k0 = bicubic_first(GR, pp);
            P acc = (P)bicubic_w(GR, pp, k0) * xv[k0];
            unroll<k0 + 1, 16>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr float w = (float)bicubic_w(GR, pp, k);
                if constexpr (w != 0.f) acc = __builtin_fmaf(w, xv[k], acc);
            });
            r[pp] = acc;
        });
        return;
    }
    unroll<0, 4>([&](auto bc) {
        constexpr int nn = 3 - decltype(bc)::value;
        P cb[4];
        unroll<0, 4>([&](auto ic) {
            constexpr int i = 4 * nn + decltype(ic)::value;
            cb[decltype(ic)::value] = wt_row<i>(xv);
        });
        unroll<0, GR * GR>([&](auto pc) {
            constexpr int pp = decltype(pc)::value, ii = pp / GR, jj = pp % GR;
            constexpr bool t2 = frac_pow2(ii, GR), u2 = frac_pow2(jj, GR);  // exact products
            const P tt = (P)(((double)ii) / ((double)GR)), uu = (P)(((double)jj) / ((double)GR));
            auto madd_u = [&](P a, P c) {  // a u + c
                if constexpr (u2) return fma_exact(a, uu, c);
                else return a * uu + c;
            };
            // X = (cb3 u + cb2) u + cb1, then X u + cb0
            auto xu_plus = [&]() { return madd_u(madd_u(madd_u(cb[3], cb[2]), cb[1]), cb[0]); };
            if constexpr (ii == 0) {
                if constexpr (nn == 0) {
                    if constexpr (jj == 0) r[pp] = cb[0];
                    else r[pp] = xu_plus();
                }
            } else if constexpr (nn == 3) {
                if constexpr (jj == 0) r[pp] = cb[0];
                else r[pp] = xu_plus();
            } else if constexpr (jj == 0) {  // t r + cb0
                if constexpr (t2) r[pp] = fma_exact(tt, r[pp], cb[0]);
                else r[pp] = tt * r[pp] + cb[0];
            } else {  // (t r + X u) + cb0
                const P X = madd_u(madd_u(cb[3], cb[2]), cb[1]);
                P m;
                if constexpr (u2) m = fma_exact(X, uu, tt * r[pp]);
                else if constexpr (t2) m = fma_exact(tt, r[pp], X * uu);
                else m = tt * r[pp] + X * uu;
                r[pp] = m + cb[0];
            }
        });
    });
}

// two adjacent values of one row, one vector store (the first at an even index)
__device__ __forceinline__ void store_pair(double *p, double a, double b) {
    *reinterpret_cast<double2 *>(p) = make_double2(a, b);
}
__device__ __forceinline__ void store_pair(float *p, float a, float b) {
    *reinterpret_cast<float2 *>(p) = make_float2(a, b);
}
__device__ __forceinline__ void store_pair(__half *p, float a, float b) {
    *reinterpret_cast<__half2 *>(p) = __floats2half2_rn(a, b);
}

// ---- fused interface fill ----------------------------------------------------------
// One workgroup = one TX x TY tile of coarse cells of one band.  For each of the
// 23 plane values in turn it
//   1. forms the Hermite-in-time value on the tile plus a 2-cell halo, in LDS
//      (the loads of the next value are issued before this value is processed);
//   2. takes d/dx, d/dy at the tile's (TX+1) x (TY+1) corners (d/dy also on the
//      halo columns), in LDS;
//   3. forms d2/dxdy = d/dx of d/dy at its cell's 4 corners, builds the cell's 16
//      bicubic coefficients and evaluates them at its GR x GR fine columns;
//   4. keeps rho and u (the first four values, each a phase of its own with a
//      fixed destination register set) in registers; for a distribution value
//      writes f = feq + tau_f/(GR tau_c) f_neq straight from registers, GR
//      adjacent fine values per store.
struct IfaceItem {
    const void *ln, *lp, *lpp;  // newest, previous, oldest plane levels [23][Yc][Xc]
    void *band;                 // band post-stream lattice
    int layer;                  // stored interface layer
    int Xc, Yc;
    int groups;                 // workgroups per tile, each filling a share of the 19 directions
    double rescale;             // tau_f / (GR tau_c)
};

template <class S, class P, int GR, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_iface(Batch<IfaceItem> BT, HermiteW<P> H) {
    if (halted(BT.halt)) return;
    constexpr int NT = TX * TY, HX = TX + 5, HY = TY + 5, NP = GR * GR;
    constexpr int NH = (HX * HY + NT - 1) / NT;  // halo points per thread
    // The value tile has two sets, used by alternate values: the next value is
    // written there before the first wait, while slower threads still read this
    // one.  The derivative tiles are written only after that wait, when every
    // thread has finished reading them, so one set serves.
    __shared__ P s0[2][HY][HX];       // value, halo 2 on each side (+1 corner)
    __shared__ P s1[TY + 1][TX + 1];  // d/dx at the corners
    __shared__ P s2[TY + 1][HX];      // d/dy at the corners and the halo columns
    const int itm = batch_find(BT.first, BT.n, blockIdx.x);
    const IfaceItem &B = BT.it[itm];
    const int lbg = blockIdx.x - BT.first[itm];
    const int G = B.groups, lb = lbg / G, grp = lbg % G;
    // this workgroup's directions [q0, q1); rho and u (values 0..3) are formed by every group
    const int q0 = (Q * grp) / G, q1 = (Q * (grp + 1)) / G, nv = 4 + q1 - q0;
    const int Xc = B.Xc, Yc = B.Yc;
    const P rescale = (P)B.rescale;
    const P *__restrict__ Ln = static_cast<const P *>(B.ln);
    const P *__restrict__ Lp = static_cast<const P *>(B.lp);
    const P *__restrict__ Lpp = static_cast<const P *>(B.lpp);
    const int ntx = (Xc + TX - 1) / TX;
    const int x0 = (lb % ntx) * TX, y0 = (lb / ntx) * TY;
    const int tid = threadIdx.x, cx = tid % TX, cy = tid / TX;
    const int n = Xc * Yc, Xf = GR * Xc, Yf = GR * Yc;
    S *band = static_cast<S *>(B.band);
    // halo points of this thread and their plane offsets (fixed for every value)
    int hoff[NH];
    bool hok[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int p = tid + k * NT;
        hok[k] = p < HX * HY;
        const int gx = ((x0 - 2 + p % HX) % Xc + Xc) % Xc, gy = ((y0 - 2 + p / HX) % Yc + Yc) % Yc;
        hoff[k] = hok[k] ? gy * Xc + gx : 0;
    }
    P a[NH][3];
#pragma unroll
    for (int k = 0; k < NH; k++)
        if (hok[k]) { a[k][0] = Ln[hoff[k]]; a[k][1] = Lp[hoff[k]]; a[k][2] = Lpp[hoff[k]]; }
    // value phase iv (value 0..3: rho, u; then directions q0..q1-1): steps 1-3 for
    // that value, its bicubic result at the cell's fine columns into r
    auto phase = [&](int iv, P *r) __attribute__((always_inline)) {
        const int sb = iv & 1;
        // 1. Hermite value on the halo tile (tinterpolate_hrmt.c:28-34)
#pragma unroll
        for (int k = 0; k < NH; k++) {
            if (!hok[k]) continue;
            const int p = tid + k * NT;
            const P ln = a[k][0], lp = a[k][1], lpp = a[k][2];
            const P c1 = lp;
            const P c2 = (ln - lpp) * (P)0.5;
            const P c3 = ln;
            const P c4 = (fma_exact((P)-4., lp, (P)3. * ln) + lpp) * (P)0.5;  // 3 ln - 4 lp
            P hv;
            if (H.exact) hv = fma_exact(H.h11, c4, fma_exact(H.h01, c3, fma_exact(H.h10, c2, H.h00 * c1)));
            else hv = H.h00 * c1 + H.h10 * c2 + H.h01 * c3 + H.h11 * c4;
            s0[sb][p / HX][p % HX] = hv;
        }
        if (iv + 1 < nv) {  // prefetch the next value
            const size_t vo = (size_t)(iv + 1 < 4 ? iv + 1 : 4 + q0 + (iv + 1 - 4)) * n;
#pragma unroll
            for (int k = 0; k < NH; k++)
                if (hok[k]) { a[k][0] = Ln[vo + hoff[k]]; a[k][1] = Lp[vo + hoff[k]]; a[k][2] = Lpp[vo + hoff[k]]; }
        }
        __syncthreads();
        // 2. d/dy on corner rows for every halo column; d/dx at the corners (derivative.c:41-42, 63-64)
        P(*t0)[HX] = s0[sb];
        for (int p = tid; p < (TY + 1) * HX; p += NT) {
            const int rr = p / HX, hx = p % HX, hy = rr + 2;
            s2[rr][hx] = d4(t0[hy - 2][hx], t0[hy - 1][hx], t0[hy + 1][hx], t0[hy + 2][hx]);
            if (hx >= 2 && hx <= TX + 2)
                s1[rr][hx - 2] = d4(t0[hy][hx - 2], t0[hy][hx - 1], t0[hy][hx + 1], t0[hy][hx + 2]);
        }
        __syncthreads();
        // 3. bicubic patch of this cell (bcucof_opt.c:47-61) at its fine columns;
        //    d2/dxdy = d/dx of d/dy (derivative.c:84) at the cell's corners, from
        //    the d/dy of columns cx .. cx+5 on its two corner rows
        P xv[16];
        P dyr[2][6];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int i = 0; i < 6; i++) dyr[j][i] = s2[cy + j][cx + i];
        xv[0] = t0[cy + 2][cx + 2]; xv[4] = s1[cy][cx];         xv[8] = dyr[0][2];
        xv[1] = t0[cy + 2][cx + 3]; xv[5] = s1[cy][cx + 1];     xv[9] = dyr[0][3];
        xv[2] = t0[cy + 3][cx + 3]; xv[6] = s1[cy + 1][cx + 1]; xv[10] = dyr[1][3];
        xv[3] = t0[cy + 3][cx + 2]; xv[7] = s1[cy + 1][cx];     xv[11] = dyr[1][2];
        xv[12] = d4(dyr[0][0], dyr[0][1], dyr[0][3], dyr[0][4]);
        xv[13] = d4(dyr[0][1], dyr[0][2], dyr[0][4], dyr[0][5]);
        xv[14] = d4(dyr[1][1], dyr[1][2], dyr[1][4], dyr[1][5]);
        xv[15] = d4(dyr[1][0], dyr[1][1], dyr[1][3], dyr[1][4]);
        bicubic_patch<GR>(xv, r);
    };
    // rho and u stay in registers for every direction of this workgroup
    P dens[NP], u0[NP], u1[NP], u2[NP];
    phase(0, dens);
    phase(1, u0);
    phase(2, u1);
    phase(3, u2);
    for (int iv = 4; iv < nv; iv++) {
        P r[NP];
        phase(iv, r);
        // 4. f = feq + tau_f/(GR tau_c) f_neq for direction q (uniform over the
        //    workgroup): one branch per direction, so e_q . u is edot<q> itself
        const int q = q0 + iv - 4;
        // feq depends on rho and u only: without this (empty) register constraint
        // the compiler hoists it out of the loop for all 19 directions at once,
        // which holds 19 NP values in registers and lowers occupancy
#pragma unroll
        for (int pp = 0; pp < NP; pp++) asm volatile("" : "+v"(dens[pp]), "+v"(u0[pp]), "+v"(u1[pp]), "+v"(u2[pp]));
        unroll<0, Q>([&](auto qc) {
            constexpr int QD = decltype(qc)::value;
            if (q != QD) return;
            const P wq = (P)WQ[QD];
            P out[NP];
#pragma unroll
            for (int pp = 0; pp < NP; pp++) {
                const P eu = edot<QD>(u0[pp], u1[pp], u2[pp]);
                // u.u formed again for every direction, the same expression each time,
                // rather than held through the loop (registers)
                const P vsq = u0[pp] * u0[pp] + u1[pp] * u1[pp] + u2[pp] * u2[pp];
                if constexpr (sizeof(S) == 8) {
                    const P feq = wq * dens[pp] * ((P)1. + (P)3. * eu + (P)4.5 * eu * eu - (P)1.5 * vsq);
                    out[pp] = feq + rescale * r[pp];  // ctfdtransfer_opt.c:93
                } else {  // dens holds rho - 1
                    const P rho = (P)1. + dens[pp];
                    const P geq = wq * (dens[pp] + rho * ((P)3. * eu + (P)4.5 * eu * eu - (P)1.5 * vsq));
                    out[pp] = geq + rescale * r[pp];
                }
            }
            // the fine columns of this cell, straight from registers: a half-wave's
            // cells cover one contiguous run of each fine row
            if (x0 + cx < Xc && y0 + cy < Yc) {
#pragma unroll
                for (int jj = 0; jj < GR; jj++) {
                    S *dst = band + lat_at(QD, (size_t)B.layer * Yf + GR * (y0 + cy) + jj, GR * (x0 + cx), Xf);
                    if constexpr (GR == 2) {
                        store_pair(dst, out[jj], out[GR + jj]);  // ii = 0, 1
                    } else {
#pragma unroll
                        for (int ii = 0; ii < GR; ii++) store_s(dst + ii, out[ii * GR + jj]);
                    }
                }
            }
        });
    }
}

// Coarse tile of one workgroup.  The 19 directions can be shared out over G
// workgroups per tile, each forming rho and u itself: that multiplies the work by
// (4 + 19/G) G / 23 but shortens each workgroup's chain of sequential value phases
// to 4 + 19/G.  G fills the workgroup slots the GPU holds at once for this kernel
// (CUs x resident workgroups per CU, from the occupancy query): a launch whose
// tiles alone fill them (the Re_tau=180 bands: 1024 tiles) takes G = 1, a launch of
// small channels (the serving queue) shortens its chains, up to G = 19.
constexpr int IF_TX = 32, IF_TY = 8;

template <class S, class P, int GR>
static int iface_slots() {
    static int slots = 0;  // per kernel; concurrent first calls compute the same value
    if (slots == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_iface<S, P, GR, IF_TX, IF_TY>, IF_TX * IF_TY, 0) ==
                hipSuccess &&
            cus > 0 && per_cu > 0)
            slots = cus * per_cu;
        else
            slots = 1024;
    }
    return slots;
}

static int iface_groups(long tiles, int slots) {
    if (tiles >= slots) return 1;
    return (int)std::min<long>(Q, slots / tiles);
}

template <class S, class P, int GR>
static hipError_t iface_launch(hipStream_t s, IfaceItem *items, int n, HermiteW<P> H, const int *halt) {
    for (int i0 = 0; i0 < n; i0 += BATCH_MAX) {
        const int m = std::min(BATCH_MAX, n - i0);
        long tiles = 0;
        for (int i = i0; i < i0 + m; i++)
            tiles += (long)((items[i].Xc + IF_TX - 1) / IF_TX) * ((items[i].Yc + IF_TY - 1) / IF_TY);
        const int G = iface_groups(tiles, iface_slots<S, P, GR>());
        for (int i = i0; i < i0 + m; i++) items[i].groups = G;
        Batch<IfaceItem> B;
        B.halt = halt;
        const int blocks = fill_batch(B, items + i0, m, [](const IfaceItem &a) {
            return ((a.Xc + IF_TX - 1) / IF_TX) * ((a.Yc + IF_TY - 1) / IF_TY) * a.groups;
        });
        hipLaunchKernelGGL((k_iface<S, P, GR, IF_TX, IF_TY>), dim3(blocks), dim3(IF_TX * IF_TY), 0, s, B, H);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <class S, class P>
static hipError_t iface_items(hipStream_t s, int GR, IfaceItem *items, int n, HermiteW<P> H, const int *halt) {
    switch (GR) {
    case 1: return iface_launch<S, P, 1>(s, items, n, H, halt);
    case 2: return iface_launch<S, P, 2>(s, items, n, H, halt);
    case 3: return iface_launch<S, P, 3>(s, items, n, H, halt);
    default: return iface_launch<S, P, 4>(s, items, n, H, halt);
    }
}

// both bands of one handle: lower writes its top stored layer, upper its layer 0
static void iface_pair(const dnsg_ctx *h, IfaceItem *out) {
    // tau_f/(GR tau_c) written with the reference's variables (omega = 1/tau):
    // tauc/((double)GR*tauf) (ctfdtransfer_opt.c:93)
    const double rescale = h->d.omega_c / ((double)h->d.GR * h->d.omega_f);
    const int zds = h->upper.nz;
    out[0] = {h->lev_lo[0], h->lev_lo[1], h->lev_lo[2], h->lower.n(), zds - 1, h->d.Xc, h->d.Yc, 1, rescale};
    out[1] = {h->lev_up[0], h->lev_up[1], h->lev_up[2], h->upper.n(), 0, h->d.Xc, h->d.Yc, 1, rescale};
}

hipError_t launch_interface_many(hipStream_t s, dnsg_ctx *const *hs, int n, int iter) {
    if (n < 1) return hipSuccess;
    const int GR = hs[0]->d.GR, prec = hs[0]->p.precision;
    // Hermite weights with the reference's expression order (tinterpolate_hrmt.c:20,34)
    const double tmp = (((double)iter) / ((double)GR));
    HermiteW<double> H;
    H.h00 = (2. * tmp * tmp * tmp - 3. * tmp * tmp + 1.);
    H.h10 = (tmp * tmp * tmp - 2. * tmp * tmp + tmp);
    H.h01 = (-2. * tmp * tmp * tmp + 3. * tmp * tmp);
    H.h11 = (tmp * tmp * tmp - tmp * tmp);
    auto pow2_or_zero = [](double w) {
        int e;
        return w == 0. || std::fabs(std::frexp(w, &e)) == 0.5;
    };
    H.exact = pow2_or_zero(H.h00) && pow2_or_zero(H.h10) && pow2_or_zero(H.h01) && pow2_or_zero(H.h11);
    // one handle (the single-solve step): its two items on the stack, so the step
    // allocates no host memory; a batch of several (the serving queue): a vector
    IfaceItem one[2];
    std::vector<IfaceItem> many(n > 1 ? 2 * (size_t)n : 0);
    IfaceItem *v = n > 1 ? many.data() : one;
    for (int i = 0; i < n; i++) iface_pair(hs[i], &v[2 * (size_t)i]);
    // one handle: its own halt flag; a batch of several (the serving queue): none
    const int *halt = n == 1 ? hs[0]->halt() : nullptr;
    if (prec == DNSG_PREC_FP64) return iface_items<double, double>(s, GR, v, 2 * n, H, halt);
    HermiteW<float> Hf = {(float)H.h00, (float)H.h10, (float)H.h01, (float)H.h11, H.exact};
    if (prec == DNSG_PREC_FP32) return iface_items<float, float>(s, GR, v, 2 * n, Hf, halt);
    return iface_items<__half, float>(s, GR, v, 2 * n, Hf, halt);
}

hipError_t launch_interface(hipStream_t s, dnsg_ctx *h, int iter) { return launch_interface_many(s, &h, 1, iter); }

// ---- fine -> coarse injection --------------------------------------------------------
struct FtcdItem {
    const void *band[2];  // lower / upper band post-stream lattice
    int layer[2];         // stored layer coincident with the coarse end layer
    int kc[2];            // coarse layer written
    void *coarse;         // coarse post-stream lattice
    int Xc, Yc, GR;
    double fac;           // GR tau_c / tau_f
    double *psum;         // if set: per workgroup [sum w e_x f, sum w f] of the values written (w = GR^3 / 2)
};

// sums of two values over a 256-thread workgroup, valid in thread 0
__device__ __forceinline__ void wg_sum2(double &a, double &b) {
    __shared__ double l[2][4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { l[0][w] = a; l[1][w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = l[0][0] + l[0][1] + l[0][2] + l[0][3];
        b = l[1][0] + l[1][1] + l[1][2] + l[1][3];
    }
}

template <class S>
__device__ __forceinline__ void ftcd_node(const FtcdItem &A, int side, int id, double &sm, double &sd) {
    const int Xc = A.Xc, Yc = A.Yc, GR = A.GR;
    const int ic = id % Xc, jc = id / Xc;
    const int Xf = GR * Xc, Yf = GR * Yc;
    const double fac = A.fac;
    const S *band = static_cast<const S *>(A.band[side]);
    S *cs = static_cast<S *>(A.coarse);
    const size_t srow = (size_t)A.layer[side] * Yf + (size_t)GR * jc, drow = (size_t)A.kc[side] * Yc + jc;
    double f[Q];
    unroll<0, Q>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        if constexpr (sizeof(S) == 8) f[q] = band[lat_at(q, srow, GR * ic, Xf)];
        else f[q] = (double)load_s(band + lat_at(q, srow, GR * ic, Xf));
    });
    if constexpr (sizeof(S) == 8) {
        const Macro64 m = macro64(f);
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const double feq = feq64<q>(m.dens, m.u0, m.u1, m.u2, m.v);
            const double fneq = f[q] - feq;
            const double out = feq + fac * fneq;
            cs[lat_at(q, drow, ic, Xc)] = out;
            sd += out;
            if constexpr (EX[q] > 0) sm += out;
            if constexpr (EX[q] < 0) sm -= out;
        });
    } else {
        const double drho = zeroth_moment(f), dens = 1. + drho;
        const double u0 = first_moment<0>(f) / dens, u1 = first_moment<1>(f) / dens, u2 = first_moment<2>(f) / dens;
        const double v = u0 * u0 + u1 * u1 + u2 * u2;
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const double eu = edot<q>(u0, u1, u2);
            const double geq = WQ[q] * (drho + dens * (3. * eu + 4.5 * eu * eu - 1.5 * v));
            const float out = (float)(geq + fac * (f[q] - geq));
            store_s(cs + lat_at(q, drow, ic, Xc), out);
            const double fv = read_back<S, q>(out);
            sd += fv;
            if constexpr (EX[q] > 0) sm += fv;
            if constexpr (EX[q] < 0) sm -= fv;
        });
    }
}

// item i: workgroups [0, nb) inject the lower band, [nb, 2 nb) the upper band
template <class S>
__global__ __launch_bounds__(256) void k_ftcd(Batch<FtcdItem> B) {
    if (halted(B.halt)) return;
    const int it = batch_find(B.first, B.n, blockIdx.x);
    const FtcdItem &A = B.it[it];
    const int Xc = A.Xc, Yc = A.Yc, GR = A.GR;
    const int n = Xc * Yc, nb = (n + 255) / 256;
    const int lb = blockIdx.x - B.first[it];
    const int side = lb >= nb ? 1 : 0;
    const int id = (lb - side * nb) * 256 + threadIdx.x;
    double sm = 0., sd = 0.;  // weighted sums of the values written (flux controller)
    if (id < n) ftcd_node<S>(A, side, id, sm, sd);
    if (A.psum) {  // uniform per workgroup
        const double w = 0.5 * GR * GR * GR;  // coarse end layers in encalc2 (encalc2.c:72-75)
        sm *= w;
        sd *= w;
        wg_sum2(sm, sd);
        if (threadIdx.x == 0) { A.psum[2 * lb] = sm; A.psum[2 * lb + 1] = sd; }
    }
}

// ---- filtered fine -> coarse injection (ftcdtransfer_lesfltr.c:16-381) ------------------
// f_neq of the fine nodes of the 5 layers kf-2 .. kf+2 around each band's coincident
// layer is box-filtered over the 5 x 5 x 5 nodes centred on every coincident node
// with the trapezoid weights w(ii) w(jj) w(kk), w = {1/2, 1, 1, 1, 1/2} (:22-28),
// each term added as (w f_neq) / GR^3 in ii, jj, kk order (:222-236); the coarse
// node gets feq(rho, m) + GR tau_c/tau_f filtered f_neq, where rho and m are the
// coincident node's unfiltered zeroth and first moments and m, the raw sum e f, is
// used as the velocity (:173-176, :207-212).  Two launches: one thread per slab
// node for f_neq (coalesced along x), one thread per coarse node for the filter.
struct FltrArgs {
    const void *band[2];    // lower / upper band post-stream lattice (stored layers)
    const double *wall[2];  // the wall layers k = 0 (lower) and k = zdf-1 (upper), [Xf][Yf][19]
    int kf[2];              // fine layer coincident with the coarse end layer
    int kc[2];              // coarse layer written
    int zdf, Xc, Yc, GR;
    double fac;             // GR tau_c / tau_f
    void *coarse;           // coarse post-stream lattice
    double *fneq;           // [side][19][5][Yf][Xf]
    double *psum;           // as FtcdItem::psum
    const int *halt;        // as Batch::halt
};

// the 19 values of fine node (i, j, k) of one band in double: f (FP64) or the
// shifted g = f - w (FP32 / FP16 storage); the band's wall layer is not stored in
// the lattice (lower: stored layer k-1, upper: stored layer k)
template <class S>
__device__ __forceinline__ void fltr_node(const FltrArgs &A, int side, int i, int j, int k, double *f) {
    const int Xf = A.GR * A.Xc, Yf = A.GR * A.Yc;
    if (side == 0 ? k == 0 : k == A.zdf - 1) {
        const double *w = A.wall[side] + ((size_t)i * Yf + j) * Q;
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            f[q] = sizeof(S) == 8 ? w[q] : w[q] - WQ[q];
        });
    } else {
        const S *b = static_cast<const S *>(A.band[side]);
        const size_t row = (size_t)(side == 0 ? k - 1 : k) * Yf + j;
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            f[q] = (double)load_s(b + lat_at(q, row, i, Xf));
        });
    }
}

template <class S>
__global__ __launch_bounds__(256) void k_fltr_slab(FltrArgs A) {
    if (halted(A.halt)) return;
    const int Xf = A.GR * A.Xc, Yf = A.GR * A.Yc;
    const size_t plane = (size_t)Xf * Yf;
    const size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= 10 * plane) return;
    const int i = (int)(id % Xf), j = (int)((id / Xf) % Yf), l = (int)(id / plane);
    const int side = l / 5, kk = l % 5;
    double f[Q];
    fltr_node<S>(A, side, i, j, A.kf[side] + kk - 2, f);
    double *o = A.fneq + ((size_t)side * Q * 5 + kk) * plane + (size_t)j * Xf + i;
    if constexpr (sizeof(S) == 8) {  // :160-200, the moments in ftcdtransfer's order
        const Macro64 m = macro64(f);
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            o[(size_t)q * 5 * plane] = f[q] - feq64<q>(m.dens, m.u0, m.u1, m.u2, m.v);
        });
    } else {
        const double drho = zeroth_moment(f), dens = 1. + drho;
        const double u0 = first_moment<0>(f) / dens, u1 = first_moment<1>(f) / dens, u2 = first_moment<2>(f) / dens;
        const double v = u0 * u0 + u1 * u1 + u2 * u2;
        unroll<0, Q>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            const double eu = edot<q>(u0, u1, u2);
            o[(size_t)q * 5 * plane] = f[q] - WQ[q] * (drho + dens * (3. * eu + 4.5 * eu * eu - 1.5 * v));
        });
    }
}

// workgroups [0, nb) filter the lower band, [nb, 2 nb) the upper band (k_ftcd's layout)
template <class S>
__global__ __launch_bounds__(256) void k_fltr(FltrArgs A) {
    if (halted(A.halt)) return;
    const int Xc = A.Xc, Yc = A.Yc, GR = A.GR, Xf = GR * Xc, Yf = GR * Yc;
    const int n = Xc * Yc, nb = (n + 255) / 256;
    const int side = (int)blockIdx.x >= nb ? 1 : 0;
    const int id = ((int)blockIdx.x - side * nb) * 256 + threadIdx.x;
    double sm = 0., sd = 0.;
    if (id < n) {
        const int ic = id % Xc, jc = id / Xc, i = GR * ic, j = GR * jc;
        const size_t plane = (size_t)Xf * Yf;
        const double *src = A.fneq + (size_t)side * Q * 5 * plane;
        constexpr double wl[5] = {0.5, 1., 1., 1., 0.5};
        const double den = (double)GR * GR * GR;
        double acc[Q];
        unroll<0, Q>([&](auto qq) { acc[decltype(qq)::value] = 0.; });
        for (int ii = 0; ii < 5; ii++) {
            const int in = (i + ii - 2 + Xf) % Xf;
            for (int jj = 0; jj < 5; jj++) {
                const int jn = (j + jj - 2 + Yf) % Yf;
                const double *p = src + (size_t)jn * Xf + in;
#pragma unroll
                for (int kk = 0; kk < 5; kk++) {
                    const double wt = wl[ii] * wl[jj] * wl[kk];
                    unroll<0, Q>([&](auto qq) {
                        constexpr int q = decltype(qq)::value;
                        acc[q] += (wt * p[((size_t)q * 5 + kk) * plane]) / den;
                    });
                }
            }
        }
        double f[Q];
        fltr_node<S>(A, side, i, j, A.kf[side], f);
        S *cs = static_cast<S *>(A.coarse);
        const size_t drow = (size_t)A.kc[side] * Yc + jc;
        // rho and the raw first moments of the coincident node (not divided by rho)
        const double dsum = zeroth_moment(f);
        const double m0 = first_moment<0>(f), m1 = first_moment<1>(f), m2 = first_moment<2>(f);
        const double v = m0 * m0 + m1 * m1 + m2 * m2;
        if constexpr (sizeof(S) == 8) {
            unroll<0, Q>([&](auto qq) {
                constexpr int q = decltype(qq)::value;
                const double out = feq64<q>(dsum, m0, m1, m2, v) + A.fac * acc[q];
                cs[lat_at(q, drow, ic, Xc)] = out;
                sd += out;
                if constexpr (EX[q] > 0) sm += out;
                if constexpr (EX[q] < 0) sm -= out;
            });
        } else {  // shifted: feq - w = w (drho + rho (3 eu + 4.5 eu^2 - 1.5 v))
            const double drho = dsum, dens = 1. + drho;
            unroll<0, Q>([&](auto qq) {
                constexpr int q = decltype(qq)::value;
                const double eu = edot<q>(m0, m1, m2);
                const double geq = WQ[q] * (drho + dens * (3. * eu + 4.5 * eu * eu - 1.5 * v));
                const float out = (float)(geq + A.fac * acc[q]);
                store_s(cs + lat_at(q, drow, ic, Xc), out);
                const double fv = read_back<S, q>(out);
                sd += fv;
                if constexpr (EX[q] > 0) sm += fv;
                if constexpr (EX[q] < 0) sm -= fv;
            });
        }
    }
    if (A.psum) {  // uniform per workgroup
        const double w = 0.5 * GR * GR * GR;
        sm *= w;
        sd *= w;
        wg_sum2(sm, sd);
        if (threadIdx.x == 0) { A.psum[2 * blockIdx.x] = sm; A.psum[2 * blockIdx.x + 1] = sd; }
    }
}

size_t fltr_fneq_doubles(const dnsg_ctx *h) { return 2 * (size_t)Q * 5 * h->d.Xf * h->d.Yf; }
size_t fltr_wall_doubles(const dnsg_ctx *h) { return 2 * (size_t)h->d.Xf * h->d.Yf * Q; }

template <class S>
static hipError_t fltr_launch(hipStream_t s, const FltrArgs &A, int nb) {
    const size_t nslab = 10 * (size_t)A.GR * A.Xc * A.GR * A.Yc;
    hipLaunchKernelGGL(k_fltr_slab<S>, dim3((unsigned)((nslab + 255) / 256)), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_fltr<S>, dim3(nb), dim3(256), 0, s, A);
    return hipGetLastError();
}

static hipError_t launch_fltr(hipStream_t s, dnsg_ctx *h, double *psum) {
    const int zdf = h->d.zdf, GR = h->d.GR;
    // the filter's index ranges: GR = 4 and a slab inside each band, 0 <= kf-2, kf+2 <= zdf-1
    if (GR != 4 || zdf - 3 - GR < 0 || !h->fneq || !h->fwall) return hipErrorInvalidValue;
    const size_t nw = (size_t)h->d.Xf * h->d.Yf * Q;
    FltrArgs A = {{h->lower.n(), h->upper.n()}, {h->fwall, h->fwall + nw}, {zdf - 1 - GR, GR}, {0, h->d.zdc - 1},
                  zdf, h->d.Xc, h->d.Yc, GR, GR * (h->d.omega_f / h->d.omega_c), h->coarse.n(), h->fneq, psum,
                  h->halt()};
    const int nb = ftcd_blocks(h);
    if (h->p.precision == DNSG_PREC_FP64) return fltr_launch<double>(s, A, nb);
    if (h->p.precision == DNSG_PREC_FP32) return fltr_launch<float>(s, A, nb);
    return fltr_launch<__half>(s, A, nb);
}

int ftcd_blocks(const dnsg_ctx *h) { return 2 * ((h->d.Xc * h->d.Yc + 255) / 256); }

hipError_t launch_ftcd_many(hipStream_t s, dnsg_ctx *const *hs, int n, double *psum) {
    if (n < 1) return hipSuccess;
    for (int i = 0; i < n; i++)  // the filtered injection runs per handle (the queue takes none)
        if (hs[i]->p.fltr) return n == 1 ? launch_fltr(s, hs[0], psum) : hipErrorInvalidValue;
    FtcdItem one[1];  // one handle: on the stack, as in launch_interface_many
    std::vector<FtcdItem> many(n > 1 ? n : 0);
    FtcdItem *v = n > 1 ? many.data() : one;
    for (int i = 0; i < n; i++) {
        const dnsg_ctx *h = hs[i];
        const int zdf = h->d.zdf, GR = h->d.GR;
        // lower band k = zdf-1-GR is stored layer zdf-2-GR; upper band k = GR is stored layer GR;
        // GR*(tauf/tauc) with omega = 1/tau (ftcdtransfer.c:66)
        v[i] = {{h->lower.n(), h->upper.n()}, {zdf - 2 - GR, GR}, {0, h->d.zdc - 1}, h->coarse.n(),
                h->d.Xc, h->d.Yc, GR, GR * (h->d.omega_f / h->d.omega_c), psum};
        if (psum) psum += 2 * (size_t)ftcd_blocks(h);
    }
    const int prec = hs[0]->p.precision;
    for (int i0 = 0; i0 < n; i0 += BATCH_MAX) {
        Batch<FtcdItem> B;
        B.halt = n == 1 ? hs[0]->halt() : nullptr;
        const int blocks = fill_batch(B, v + i0, std::min(BATCH_MAX, n - i0),
                                      [](const FtcdItem &a) { return 2 * ((a.Xc * a.Yc + 255) / 256); });
        if (prec == DNSG_PREC_FP64)
            hipLaunchKernelGGL(k_ftcd<double>, dim3(blocks), dim3(256), 0, s, B);
        else if (prec == DNSG_PREC_FP32)
            hipLaunchKernelGGL(k_ftcd<float>, dim3(blocks), dim3(256), 0, s, B);
        else
            hipLaunchKernelGGL(k_ftcd<__half>, dim3(blocks), dim3(256), 0, s, B);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_ftcd(hipStream_t s, dnsg_ctx *h, double *psum) { return launch_ftcd_many(s, &h, 1, psum); }

}  // namespace dnsg
