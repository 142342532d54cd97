// spectra.hip -- 1-D energy spectra of the velocity planes (dnsg_spectra).
//
// The reference computes them offline from the velocity files (e-spectra/src/
// 1d-spectra-mgrd.c:119-230): per wall-normal plane a 2-D real-to-complex DFT over
// (x streamwise, y spanwise) of u, v, w, scaled by 1/(N0 N1), the mean mode
// removed, then
//   E_x(i) = sum_j fac_j 1/2 |c(i, j)|^2   (streamwise wavenumber index i, all N0)
//   E_y(j) = sum_i fac_j 1/2 |c(i, j)|^2   (spanwise index j = 0 .. N1/2)
// with fac_j = 1 for j = 0 and 2 otherwise (the folded half of the spectrum).
// Here the velocity planes are formed on the device from the lattice (the
// vel_print.c formula), transformed with hipFFT in one batched plan, and folded
// by one thread per output value in the reference's summation order.
#include <hipfft/hipfft.h>

#include <vector>

#include "internal.hpp"
#include "lattice.hpp"

namespace dnsg {

namespace {

// velocity of every node of nz_st stored layers, into planes k0.. of
// vel[3][nzg][nx][ny] (y fastest: the DFT's last, halved dimension)
template <class S>
__global__ __launch_bounds__(256) void k_vel_planes(const S *__restrict__ lat, int nx, int ny, int nz_st, int k0,
                                                    int nzg, double *__restrict__ vel) {
    const size_t n = (size_t)nx * ny * nz_st;
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int x = (int)(id % nx), y = (int)((id / nx) % ny), m = (int)(id / ((size_t)nx * ny));
    double ux = 0., uy = 0., uz = 0., dens = 0.;
    unroll<0, Q>([&](auto qc) {  // vel_print.c:86-104, in its order
        constexpr int q = decltype(qc)::value;
        double f = (double)load_s(lat + lat_at(q, (size_t)m * ny + y, x, nx));
        if constexpr (sizeof(S) != 8) f += WQ[q];
        ux += (double)EX[q] * f;
        uy += (double)EY[q] * f;
        uz += (double)EZ[q] * f;
        dens += f;
    });
    if (dens > 0.) { ux /= dens; uy /= dens; uz /= dens; }
    else { ux = 0.; uy = 0.; uz = 0.; }
    const size_t plane = (size_t)nx * ny, at = (size_t)x * ny + y, k = (size_t)(m + k0);
    vel[(0 * (size_t)nzg + k) * plane + at] = ux;
    vel[(1 * (size_t)nzg + k) * plane + at] = uy;
    vel[(2 * (size_t)nzg + k) * plane + at] = uz;
}

// one output value per thread: E_x(c, i, k) for t < 3 nzg N0, then E_y(c, j, k)
__global__ __launch_bounds__(256) void k_fold(const double2 *__restrict__ cf, int N0, int N1, int nzg,
                                              double *__restrict__ ex, double *__restrict__ ey) {
    const int N1h = N1 / 2 + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nxo = 3L * nzg * N0, nyo = 3L * nzg * N1h;
    if (t >= nxo + nyo) return;
    const double s = 2. * (1. / ((double)(2. * N0 * N1)));  // 1d-spectra-mgrd.c:183
    auto term = [&](const double2 *p, int i, int j) {
        if (i == 0 && j == 0) return 0.;  // the mean removed (:195-200)
        const double2 c = p[(size_t)i * N1h + j];
        const double re = c.x * s, im = c.y * s;
        const double fac = j == 0 ? 1. : 2.;
        return fac * 0.5 * (re * re + im * im);
    };
    if (t < nxo) {
        const int i = (int)(t % N0), k = (int)((t / N0) % nzg), c = (int)(t / ((long)N0 * nzg));
        const double2 *p = cf + ((size_t)c * nzg + k) * (size_t)N0 * N1h;
        double e = 0.;
        for (int j = 0; j < N1h; j++) e += term(p, i, j);
        ex[((size_t)c * N0 + i) * nzg + k] = e;
    } else {
        const long u = t - nxo;
        const int j = (int)(u % N1h), k = (int)((u / N1h) % nzg), c = (int)(u / ((long)N1h * nzg));
        const double2 *p = cf + ((size_t)c * nzg + k) * (size_t)N0 * N1h;
        double e = 0.;
        for (int i = 0; i < N0; i++) e += term(p, i, j);
        ey[((size_t)c * N1h + j) * nzg + k] = e;
    }
}

// vel_print's velocity of one host-kept wall layer [X][Y][19] -> 3 planes [x][y]
void wall_velocity(const double *w, int nx, int ny, std::vector<double> &out) {
    static const int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0};
    static const int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
    static const int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    const size_t plane = (size_t)nx * ny;
    out.assign(3 * plane, 0.);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const double *f = w + ((size_t)i * ny + j) * 19;
            double ux = 0., uy = 0., uz = 0., dens = 0.;
            for (int a = 0; a < 19; a++) {
                ux += ex[a] * f[a];
                uy += ey[a] * f[a];
                uz += ez[a] * f[a];
                dens += f[a];
            }
            if (dens > 0.) { ux /= dens; uy /= dens; uz /= dens; }
            else { ux = 0.; uy = 0.; uz = 0.; }
            out[0 * plane + (size_t)i * ny + j] = ux;
            out[1 * plane + (size_t)i * ny + j] = uy;
            out[2 * plane + (size_t)i * ny + j] = uz;
        }
}

}  // namespace

}  // namespace dnsg

using namespace dnsg;

#define SPCHK(h, expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            rc = e_ == hipErrorOutOfMemory ? DNSG_ERR_OOM : DNSG_ERR_HIP;                   \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)

extern "C" int dnsg_spectra(dnsg_t *h, int grid, double *ex, double *ey) {
    if (!h || !h->stream || grid < 0 || grid > 2 || !ex || !ey) return DNSG_ERR_ARG;
    if (!h->have_state) {
        h->err = "no state: call dnsg_init_state or dnsg_upload_state first";
        return DNSG_ERR_STATE;
    }
    const Grid &g = grid == 0 ? h->coarse : (grid == 1 ? h->lower : h->upper);
    const int N0 = g.nx, N1 = g.ny, N1h = N1 / 2 + 1;
    const int nzg = grid == 0 ? h->d.zdc : h->d.zdf;  // planes in the velocity file
    const int k0 = grid == 1 ? 1 : 0;                 // lower band: the wall is plane 0
    const size_t plane = (size_t)N0 * N1;
    int rc = DNSG_OK;
    double *vel = nullptr, *dex = nullptr, *dey = nullptr;
    double2 *cf = nullptr;
    hipfftHandle plan = 0;
    bool have_plan = false;
    SPCHK(h, hipMalloc(&vel, 3 * (size_t)nzg * plane * sizeof(double)));
    SPCHK(h, hipMalloc(&cf, 3 * (size_t)nzg * N0 * N1h * sizeof(double2)));
    SPCHK(h, hipMalloc(&dex, 3 * (size_t)N0 * nzg * sizeof(double)));
    SPCHK(h, hipMalloc(&dey, 3 * (size_t)N1h * nzg * sizeof(double)));
    {
        const size_t n = plane * g.nz;
        const dim3 grd((unsigned)((n + 255) / 256)), blk(256);
        if (h->p.precision == DNSG_PREC_FP64)
            hipLaunchKernelGGL(k_vel_planes<double>, grd, blk, 0, h->stream, (const double *)g.c(), N0, N1, g.nz, k0, nzg, vel);
        else if (h->p.precision == DNSG_PREC_FP32)
            hipLaunchKernelGGL(k_vel_planes<float>, grd, blk, 0, h->stream, (const float *)g.c(), N0, N1, g.nz, k0, nzg, vel);
        else
            hipLaunchKernelGGL(k_vel_planes<__half>, grd, blk, 0, h->stream, (const __half *)g.c(), N0, N1, g.nz, k0, nzg, vel);
        SPCHK(h, hipGetLastError());
    }
    if (grid != 0) {  // the band's wall layer is kept on the host
        std::vector<double> wv;
        wall_velocity(grid == 1 ? h->host_wall_lo : h->host_wall_up, N0, N1, wv);
        const int kw = grid == 1 ? 0 : nzg - 1;
        for (int c = 0; c < 3; c++)
            SPCHK(h, hipMemcpyAsync(vel + ((size_t)c * nzg + kw) * plane, wv.data() + c * plane, plane * sizeof(double),
                                    hipMemcpyHostToDevice, h->stream));
        SPCHK(h, hipStreamSynchronize(h->stream));  // wv goes out of scope
    }
    {
        int dims[2] = {N0, N1};
        if (hipfftPlanMany(&plan, 2, dims, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_D2Z, 3 * nzg) != HIPFFT_SUCCESS) {
            h->err = "hipfftPlanMany failed";
            rc = DNSG_ERR_HIP;
            goto done;
        }
        have_plan = true;
        if (hipfftSetStream(plan, h->stream) != HIPFFT_SUCCESS ||
            hipfftExecD2Z(plan, vel, reinterpret_cast<hipfftDoubleComplex *>(cf)) != HIPFFT_SUCCESS) {
            h->err = "hipfftExecD2Z failed";
            rc = DNSG_ERR_HIP;
            goto done;
        }
    }
    {
        const long nout = 3L * nzg * (N0 + N1h);
        hipLaunchKernelGGL(k_fold, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, h->stream, (const double2 *)cf, N0,
                           N1, nzg, dex, dey);
        SPCHK(h, hipGetLastError());
    }
    SPCHK(h, hipMemcpyAsync(ex, dex, 3 * (size_t)N0 * nzg * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SPCHK(h, hipMemcpyAsync(ey, dey, 3 * (size_t)N1h * nzg * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SPCHK(h, hipStreamSynchronize(h->stream));
done:
    if (rc) (void)hipStreamSynchronize(h->stream);
    if (have_plan) hipfftDestroy(plan);
    if (vel) (void)hipFree(vel);
    if (cf) (void)hipFree(cf);
    if (dex) (void)hipFree(dex);
    if (dey) (void)hipFree(dey);
    return rc;
}
