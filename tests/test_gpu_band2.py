"""The two-substep band kernel (k_band2.hip) in the shifted tiers against one
collide-stream launch per substep, through the C-ABI.  DNSG_BAND2=1 / 0 at handle
creation selects either path; FP32 and FP16 run the kernel with 8-row tiles (FP64,
4 rows, is covered against the oracle by test_gpu_parity.py).

Both paths collide every node with the same arithmetic and write the same values,
so under a constant body force the states are bitwise equal after any number of
steps.  The fused reductions sum the same values in other partial layouts (64-column
tiles against 256-node chunks), so flux and density agree to 1e-12 and the wall
stresses to 1e-10 (relative).  Under the flux controller the force feeds back from
those sums: the states agree to the tier's rounding (1e-5 in FP32, 1e-2 in FP16
storage).  Cases: rows of 512 and 1024 nodes with 16-row bands (two full tiles);
rows of 80 (a partial 16-column tile) with 12-row bands (a tile of 8 rows and one of
4); stripes and posts."""
import numpy as np
import pytest

from helpers import dnsg, relmax

pytestmark = pytest.mark.gpu

BASE = dict(zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
            init_kind="laminar_noise", seed=11, u_max=0.05, noise=0.05)
CASES = {
    "x512_rslip": dict(BASE, xdv=512, ydv=8, pattern="rslip", gyplussize=4, wyplussize=4),
    "x256_pslip": dict(BASE, xdv=256, ydv=8, pattern="pslip", gxplussize=8, wxplussize=8, gyplussize=4,
                       wyplussize=4),
    "x40_y6_pslip": dict(BASE, xdv=40, ydv=6, pattern="pslip", gxplussize=6, wxplussize=2, gyplussize=3,
                         wyplussize=3),
}
GRIDS = ("coarse", "lower", "upper")


def run(kw, prec, band2, n, monkeypatch, **extra):
    monkeypatch.setenv("DNSG_BAND2", "1" if band2 else "0")
    s = dnsg.Solver(dnsg.make_params(**kw, precision=prec, **extra))
    assert s.band_launch_substeps() == (2 if band2 else 1)
    s.init_state()
    dg = s.step(n)
    st = s.download()
    s.close()
    return st, dg


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
@pytest.mark.parametrize("name", list(CASES))
def test_band2_equals_two_launches_constant_force(name, prec, monkeypatch):
    kw = CASES[name]
    a, da = run(kw, prec, True, 12, monkeypatch, force_mode="const", g_const=2e-5)
    b, db = run(kw, prec, False, 12, monkeypatch, force_mode="const", g_const=2e-5)
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(a[k], b[k]), f"{k}: max diff {np.max(np.abs(a[k] - b[k]))}"
    for f in ("flowrate", "density"):
        assert relmax(da[f], db[f]) <= 1e-12, f
    for f in ("str_up", "str_down", "stress"):
        assert relmax(da[f], db[f]) <= 1e-10, f


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_band2_equals_two_launches_flux_control(prec, monkeypatch):
    kw = CASES["x40_y6_pslip"]
    a, da = run(kw, prec, True, 20, monkeypatch)
    b, db = run(kw, prec, False, 20, monkeypatch)
    tol = {"fp32": 1e-5, "fp16": 1e-2}[prec]
    for k in GRIDS:
        assert relmax(a[k], b[k]) <= tol, (k, relmax(a[k], b[k]))
    assert relmax(da["flowrate"], db["flowrate"]) <= tol
    assert np.max(np.abs(da["force"] - db["force"])) <= tol * kw["ub"]


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_band2_nan_stops_where_two_launches_stop(prec, monkeypatch):
    """A NaN planted in the middle of the coarse grid reaches the bands' wall layers a
    few steps later.  Asked for many more steps in one call, both band paths return
    the NaN status at the same step, with the same records up to it, and hold the same
    state: that step's top (the band pass swaps the band buffers once per step, the
    two-launch path twice, and each rewinds by its own count)."""
    kw = CASES["x40_y6_pslip"]
    out = {}
    for band2 in (True, False):
        monkeypatch.setenv("DNSG_BAND2", "1" if band2 else "0")
        s = dnsg.Solver(dnsg.make_params(**kw, precision=prec, force_mode="const", g_const=2e-5))
        s.init_state()
        st = s.download()
        st["coarse"][20, 3, 6, 0] = np.nan
        s.upload(st["coarse"], st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
        n = 60
        d = (dnsg.Diag * n)()
        rc = s.lib.dnsg_step(s.h, n, d)
        msg = s.lib.dnsg_last_error(s.h).decode()
        print(f"{prec} band2={band2}: rc {rc}, {msg}", flush=True)
        assert rc == -3
        fail = int(msg.rsplit(" ", 1)[1])
        out[band2] = (fail, [(d[k].time, d[k].force, d[k].stress) for k in range(fail + 1)], s.download())
        s.close()
    (fa, ra, ga), (fb, rb, gb) = out[True], out[False]
    assert fa == fb and fa >= 1
    for (ta, fca, sa), (tb, fcb, sb) in zip(ra[:-1], rb[:-1]):
        assert ta == tb and fca == fcb
        assert relmax(np.array([sa]), np.array([sb])) <= 1e-10
    assert np.isnan(ra[-1][2]) and np.isnan(rb[-1][2])
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(ga[k], gb[k], equal_nan=True), k
