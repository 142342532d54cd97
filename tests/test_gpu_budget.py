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
import os
import subprocess

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


POSTPROC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                        "postproc_micro", "postprocess-budget")


def _dat_col2(path):
    vals = []
    for line in open(path):
        parts = line.split()
        try:
            vals.append([float(x) for x in parts])
        except ValueError:
            continue
    return np.array([v[1] for v in vals if len(v) == 2])


def test_time_averages_against_reference_postprocessor(tmp_path):
    """dnsg_budget2d_accumulate / _average over 4 snapshots of the 16x8x12 stripe channel
    against the reference's own post-processor (TKE-budget/postprocessing, built from
    its sources by oracle/build_postproc_ref.sh) run on the same snapshots' files: its
    wall-unit profiles Uavg.dat (mean U), utin.dat (rms of u'), omega-rms-x-pl.dat
    (rms vorticity), rests.dat (-u'w') and Pii.dat (production), recomputed from the
    library's averages with u_tau = sqrt(<p-grad>) over the same window as the tool
    (time_stat.c:85-100), agree to the tool's 9 printed decimals.  The library folds
    each snapshot's profiles (mean-of-squares for the rms rows, as the tool does);
    the average equals the host mean of the returned per-snapshot profiles."""
    if not os.path.exists(POSTPROC):
        pytest.skip("reference post-processor not built (oracle/build_postproc_ref.sh needs the reference sources)")
    kw = dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0, pattern="rslip",
              gyplussize=4, wyplussize=4, force_mode="const", g_const=2e-5, init_kind="laminar_noise", seed=42,
              u_max=0.05, noise=0.05)
    s = dnsg.Solver(dnsg.make_params(**kw))
    s.init_state()
    d = s.dims
    NZ, ut = d.NZ, d.utime
    recs = []
    snaps = []
    for k in range(4):
        if k:
            recs.append(s.step(ut))
        t, v, e = s.budget2d_accumulate()
        snaps.append((t, v, e))
        for name, arr in (("turb-fld-2d", t), ("vort-fluc-2d", v), ("KE-Budget-2d", e)):
            arr.tofile(tmp_path / f"{name}.000.{k:04d}")
    recs.append(s.step(1))  # the record at the last snapshot's step
    avg_t, avg_v, avg_e, cnt = s.budget2d_average()
    assert cnt == 4
    sq_t = [3, 4, 5, 8]
    host_t = np.mean([np.where(np.isin(np.arange(19), sq_t)[:, None], x[0] ** 2, x[0]) for x in snaps], axis=0)
    assert np.allclose(avg_t, host_t, rtol=1e-15, atol=0)
    assert np.allclose(avg_v, np.mean([x[1] ** 2 for x in snaps], axis=0), rtol=1e-15, atol=0)
    assert np.allclose(avg_e, np.mean([x[2] for x in snaps], axis=0), rtol=1e-15, atol=0)
    for log, key in (("p-grad.txt", "pgrad"), ("flowrate.txt", "flowrate")):
        with open(tmp_path / log, "w") as f:
            for r in recs:
                for tt, val in zip(r["time"], r[key]):
                    f.write(f"{tt:f} {val:.12f}\n")
    rp = subprocess.run([POSTPROC], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert rp.returncode == 0, rp.stdout[-1000:] + rp.stderr[-1000:]
    # u_tau as time_stat.c:85-100 forms it: the mean p-grad of the log lines in [TSTR, TEND]
    lines = np.loadtxt(tmp_path / "p-grad.txt", ndmin=2)
    times = np.array([float(f"{t:f}") for t in lines[:, 0]])
    utau = np.sqrt(lines[(times >= 0) & (times <= 3), 1].mean())
    tau = (3. / 20.0) * 0.0333333333333 * (d.NZs - 2.) + 0.5
    nu = (1. / 6.) * (2. * tau - 1.)
    up = 1 + (NZ - 1) // 2
    k = np.arange(up)
    fold = lambda a: 0.5 * (a[k] + a[NZ - 1 - k])
    expect = {
        "Uavg.dat": fold(avg_t[0]) / utau,
        "utin.dat": np.sqrt(fold(avg_t[3])) / utau,
        "rests.dat": 0.5 * (-avg_t[6][k] + avg_t[6][NZ - 1 - k]) / utau ** 2,
        "omega-rms-x-pl.dat": np.sqrt(fold(avg_v[0])) * nu / utau ** 2,
        "Pii.dat": fold(avg_e[0]) * nu / utau ** 4,
    }
    for name, want in expect.items():
        got = _dat_col2(tmp_path / name)
        print(f"time average {name}: max |tool - library| {np.max(np.abs(got - want)):.2e}", flush=True)
        assert got.shape == want.shape, name
        assert np.all(np.abs(got - want) <= 2e-9 + 1e-8 * np.abs(want)), name
