"""The paired collide-stream form of the shifted tiers (two adjacent nodes per
thread, 2-value loads and stores) against the one-node scatter form, through the
C-ABI.  The form is chosen per lattice where a row holds a multiple of 512 nodes;
DNSG_CS_PAIRED=0 at handle creation selects the scatter form everywhere.

Both forms collide every node with the same arithmetic and write the same set of
values, so with a constant body force the states are bitwise equal after any
number of steps; the fused reductions sum the same values in another order, so the
reported flux, density and stresses agree to 1e-12 (relative).  Under the flux
controller the force feeds back from those sums: the states then agree to the
tier's rounding (1e-12 relative in FP64, whose launches pair up only when they sum
their writes for the controller, 1e-5 in FP32, 1e-2 in FP16 storage) after the run.
Cases: coarse rows of 512 (paired), bands of 1024 (paired) and a lattice whose
coarse rows of 256 stay in the scatter form while its bands of 512 pair up; walls
with stripes and posts."""
import numpy as np
import pytest

from helpers import dnsg, relmax

pytestmark = pytest.mark.gpu

BASE = dict(ydv=8, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
            init_kind="laminar_noise", seed=5, u_max=0.05, noise=0.05)
CASES = {
    "x512_rslip": dict(BASE, xdv=512, pattern="rslip", gyplussize=4, wyplussize=4),
    "x512_pslip": dict(BASE, xdv=512, pattern="pslip", gxplussize=8, wxplussize=8, gyplussize=4, wyplussize=4),
    "x256_pslip": dict(BASE, xdv=256, pattern="pslip", gxplussize=6, wxplussize=2, gyplussize=4, wyplussize=4),
}
GRIDS = ("coarse", "lower", "upper")


def run(kw, prec, paired, n, monkeypatch, **extra):
    monkeypatch.setenv("DNSG_CS_PAIRED", "1" if paired else "0")
    s = dnsg.Solver(dnsg.make_params(**kw, precision=prec, **extra))
    s.init_state()
    dg = s.step(n)
    st = s.download()
    s.close()
    return st, dg


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
@pytest.mark.parametrize("name", list(CASES))
def test_paired_equals_scatter_constant_force(name, prec, monkeypatch):
    kw = CASES[name]
    a, da = run(kw, prec, True, 12, monkeypatch, force_mode="const", g_const=2e-5)
    b, db = run(kw, prec, False, 12, monkeypatch, force_mode="const", g_const=2e-5)
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(a[k], b[k]), f"{k}: max diff {np.max(np.abs(a[k] - b[k]))}"
    for f in ("flowrate", "density"):
        assert relmax(da[f], db[f]) <= 1e-12, f
    for f in ("str_up", "str_down", "stress"):
        assert relmax(da[f], db[f]) <= 1e-10, f


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16"])
def test_paired_equals_scatter_flux_control(prec, monkeypatch):
    """FP64 pairs only in launches that sum their writes, which flux control asks for"""
    kw = CASES["x512_pslip"]
    a, da = run(kw, prec, True, 20, monkeypatch)
    b, db = run(kw, prec, False, 20, monkeypatch)
    tol = {"fp64": 1e-12, "fp32": 1e-5, "fp16": 1e-2}[prec]
    for k in GRIDS:
        assert relmax(a[k], b[k]) <= tol, (k, relmax(a[k], b[k]))
    assert relmax(da["flowrate"], db["flowrate"]) <= tol
    assert np.max(np.abs(da["force"] - db["force"])) <= tol * kw["ub"]
