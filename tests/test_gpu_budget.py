"""The TKE-budget tool's per-snapshot pass on the device (dnsg_budget2d) against the
oracle's restatement of it (oracle/budget_oracle.c, itself pinned byte for byte to
the tool built from its own sources, tests/test_oracle_golden.py) on the same
snapshot.

Each node's terms are formed as the tool forms them; the layer sums (the means
included) are taken in another order, so the profiles agree to 1e-10 of their
scale (helpers.budget_scales: a row's own magnitude, or for sums of signed
products that can cancel, the magnitude of what they are made of), and the
fluctuation planes to 1e-12.
"""
import numpy as np
import pytest

from helpers import budget_scales, dnsg, oracle_budget

pytestmark = pytest.mark.gpu

INIT = dict(init_kind="laminar_noise", seed=42, u_max=0.05, noise=0.1)
CASES = {
    # the thinnest bands: the one-sided z stencils reach into the neighbouring rows
    "micro": dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
                  pattern="rslip", gyplussize=4, wyplussize=4),
    # grid ratio 3, deeper bands, posts
    "gr3": dict(xdv=12, ydv=8, zdv=14, gratio=3, zds=8, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
                pattern="pslip", gxplussize=6, wxplussize=6, gyplussize=6, wyplussize=6),
}
NAMES = {"turb": ["UAVG", "VAVG", "WAVG", "UTIN", "VTIN", "WTIN", "uw", "DAVG", "DTIN", "SKEWu", "SKEWv", "SKEWw",
                  "SKEWuw", "SKEWp", "KURTu", "KURTv", "KURTw", "KURTuw", "KURTp"],
         "vort": ["wx", "wy", "wz"],
         "ke": ["Pii", "Puu", "Pvv", "Pww", "Puw", "Tiir", "Tuur", "Tvvr", "Twwr", "Tuwr", "Tiipai", "Tuupai",
                "Tvvpai", "Twwpai", "Tuwpai", "Paiuu", "Paivv", "Paiww", "Paiuw", "Tiis", "Tuus", "Tvvs", "Twws",
                "Tuws", "Epii", "Epuu", "Epvv", "Epww", "Epuw", "KE"]}


def _close(got, ref, kind, tol):
    scale = budget_scales(kind, ref)
    for r, name in enumerate(NAMES[kind]):
        err = np.max(np.abs(got[r] - ref[r]))
        assert err <= tol * scale[r] + 1e-300, f"{name}: {err:.3e} against scale {scale[r]:.3e}"


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(CASES))
def test_budget_matches_oracle(name, prec):
    p = dnsg.make_params(**CASES[name], **INIT, force_mode="const", g_const=2e-5, precision=prec)
    s = dnsg.Solver(p)
    s.init_state()
    s.step(6)
    print(f"budget {name} {prec}", flush=True)
    turb, vort, ke, (ufl, vfl) = s.budget2d(planes=True)
    snaps = {g: s.velocity(g) for g in (0, 1, 2)}
    s.close()
    rt, rv, rk, (rufl, rvfl) = oracle_budget(p, snaps, planes=True)
    _close(turb, rt, "turb", 1e-10)
    _close(vort, rv, "vort", 1e-10)
    _close(ke, rk, "ke", 1e-10)
    assert np.max(np.abs(ufl - rufl)) <= 1e-12 * np.max(np.abs(rufl))
    assert np.max(np.abs(vfl - rvfl)) <= 1e-12 * np.max(np.abs(rvfl))
    # the statistics are not trivial: fluctuations, vorticity and a budget exist
    assert np.all(turb[3, 1:-1] > 0) and np.all(vort[:, 1:-1] > 0) and np.max(np.abs(ke[0])) > 0


def test_budget_of_a_laminar_state():
    """the seeded profile without noise: no fluctuations, so every fluctuation moment,
    vorticity fluctuation and budget term is zero, and UAVG is the laminar profile"""
    p = dnsg.make_params(**CASES["micro"], init_kind="laminar_noise", seed=1, u_max=0.05, noise=0.0)
    s = dnsg.Solver(p)
    s.init_state()
    turb, vort, ke, _ = s.budget2d()
    s.close()
    assert np.max(np.abs(turb[[3, 4, 5, 6]])) <= 1e-15  # rms u, v, w and uw
    assert np.max(np.abs(vort)) <= 1e-15
    assert np.max(np.abs(ke[:-1][[0, 1, 4, 5, 6, 9, 24]])) <= 1e-20
    d = dnsg.dims_of(p)
    z = np.zeros(d.NZ)  # wall distances in fine units (statistics.c:320-325)
    z[1:d.zdf] = np.arange(1, d.zdf) - 0.5
    k = np.arange(2, d.zdc - 2)
    z[d.zdf - 2 + k] = d.GR * (k - 1) + (d.zdf - 1 - 0.5)
    k = np.arange(0, d.zdf - 1)
    z[d.zdf + d.zdc - 4 + k] = k + d.GR * (d.zdc - 3) + d.zdf - 1 - 0.5
    zeta = z / (d.NZs - 2)
    lam = 4 * 0.05 * zeta * (1 - zeta)
    inner = np.r_[1:d.zdf, d.zdf:d.zdf + d.zdc - 4, d.zdf + d.zdc - 4:d.NZ - 1]
    assert np.max(np.abs(turb[0, inner] - lam[inner])) <= 1e-15
