"""On-device 1-D energy spectra (dnsg_spectra) against the definition of the
reference's offline tool (e-spectra/src/1d-spectra-mgrd.c:119-230).

* numpy's FFT of the velocity planes (as vel_print writes them, taken from the
  same state) folded with the reference's scaling, mean removal and R2C factor
  gives the same spectra to 1e-12 (FFTW itself is not available here: numpy's
  transform stands in for it as the mathematical definition);
* a single streamwise Fourier mode u = A cos(2 pi m x / N0) puts A^2/8 at the
  streamwise indices m and N0 - m, A^2/4 at spanwise index 0, nothing elsewhere;
* Parseval: both spectra sum to half the plane's mean square fluctuation plus
  the spanwise Nyquist column once more (the reference folds it twice).
"""
import numpy as np
import pytest

from helpers import Oracle, dnsg

pytestmark = pytest.mark.gpu

KW = dict(xdv=32, ydv=16, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0, pattern="rslip",
          gyplussize=4, wyplussize=4, force_mode="const", g_const=2e-5, init_kind="laminar_noise", seed=11,
          u_max=0.05, noise=0.2)


def reference_spectra(vel):
    """vel [N0][N1][nz][3] -> (ex [3][N0][nz], ey [3][N1/2+1][nz]) by the reference's formulas"""
    N0, N1, nz, _ = vel.shape
    c = np.fft.rfft2(np.moveaxis(vel, (2, 3), (1, 0)), axes=(2, 3))  # [3][nz][N0][N1/2+1]
    c = c * (2.0 * (1.0 / (2.0 * N0 * N1)))
    c[:, :, 0, 0] = 0.0
    fac = np.ones(N1 // 2 + 1)
    fac[1:] = 2.0
    e = fac * 0.5 * (c.real ** 2 + c.imag ** 2)  # [3][nz][N0][N1h]
    ex = np.moveaxis(e.sum(axis=3), 1, 2)        # [3][N0][nz]
    ey = np.moveaxis(e.sum(axis=2), 1, 2)        # [3][N1h][nz]
    return ex, ey


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spectra_match_definition(prec):
    s = dnsg.Solver(dnsg.make_params(**KW, precision=prec))
    s.init_state()
    s.step(20)
    for grid in (0, 1, 2):
        vel, _ = s.velocity(grid)
        ex, ey = s.spectra(grid)
        rx, ry = reference_spectra(vel)
        assert np.max(np.abs(ex - rx)) <= 1e-12 * np.max(np.abs(rx)), grid
        assert np.max(np.abs(ey - ry)) <= 1e-12 * np.max(np.abs(ry)), grid
        # Parseval: the sums over wavenumbers are half the plane's mean square
        # fluctuation, plus one more copy of the spanwise Nyquist column: the
        # reference doubles every j >= 1, j = N1/2 included (1d-spectra-mgrd.c:206-210)
        u = np.moveaxis(vel, 3, 0)  # [3][N0][N1][nz]
        N0, N1 = u.shape[1], u.shape[2]
        half_ms = 0.5 * ((u - u.mean(axis=(1, 2), keepdims=True)) ** 2).mean(axis=(1, 2))
        c = np.fft.rfft2(np.moveaxis(u, 3, 1), axes=(2, 3)) / (N0 * N1)  # [3][nz][N0][N1h]
        nyq = 0.5 * (np.abs(c[:, :, :, N1 // 2]) ** 2).sum(axis=2) if N1 % 2 == 0 else 0.0
        assert np.allclose(ex.sum(axis=1), half_ms + nyq, rtol=1e-10, atol=1e-18)
        assert np.allclose(ey.sum(axis=1), half_ms + nyq, rtol=1e-10, atol=1e-18)
    s.close()


def test_single_mode_closed_form():
    p = dnsg.make_params(**KW)
    o = Oracle(p)
    o.init_seeded()
    st = o.state()
    o.close()
    d = dnsg.dims_of(p)
    N0, A, m = d.Xc, 0.01, 3
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    e = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0],
                  [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1],
                  [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]], dtype=float)
    ux = A * np.cos(2 * np.pi * m * np.arange(N0) / N0)  # along the coarse x
    eu = e[0][None, :] * ux[:, None]
    feq = w[None, :] * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (ux ** 2)[:, None])  # rho = 1, u = (ux, 0, 0)
    coarse = np.broadcast_to(feq[:, None, None, :], st["coarse"].shape).copy()
    s = dnsg.Solver(p)
    s.upload(coarse, st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
    ex, ey = s.spectra(0)
    s.close()
    vel_x = np.zeros(N0)
    vel_x[[m, N0 - m]] = A * A / 8
    for k in range(d.zdc):
        assert np.allclose(ex[0, :, k], vel_x, rtol=1e-9, atol=1e-20), k
        assert abs(ey[0, 0, k] - A * A / 4) <= 1e-9 * A * A
        assert np.all(np.abs(ey[0, 1:, k]) <= 1e-20)
    assert np.all(np.abs(ex[1:]) <= 1e-20) and np.all(np.abs(ey[1:]) <= 1e-20)
