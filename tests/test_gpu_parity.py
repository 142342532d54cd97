"""Parity of the HIP path (through the C-ABI) with the oracle.

Bar (FP64 plain path):
  * constant body force: the state after every step is bitwise equal to the
    oracle's (no reduction feeds back into the state);
  * flux-controlled force (the reference's mode): G = ub - <rho u> comes from a
    parallel reduction whose summation order differs from the reference's
    sequential loop, so fields agree to 1e-12 relative max-norm after the run
    (tier (a): 200 steps of the 64x32x32 channel);
  * reductions (flux, density, wall stress, mass, statistics): 1e-12 relative
    max-norm per time series / profile (stresses: 1e-10, they are differences
    of nearly equal velocities).
The oracle itself is pinned bit-for-bit to the reference by
tests/test_oracle_golden.py.
"""
import numpy as np
import pytest

from helpers import Oracle, dnsg, relmax, uniform_field

pytestmark = pytest.mark.gpu

MICRO = dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0)
CASES = {
    "none": dict(MICRO),
    "rslip": dict(MICRO, pattern="rslip", gyplussize=4, wyplussize=4),
    "pslip": dict(MICRO, pattern="pslip", gxplussize=4, wxplussize=4, gyplussize=4, wyplussize=4),
    "sslip": dict(MICRO, pattern="sslip", gxplussize=6, wxplussize=2),
    "gr3": dict(MICRO, gratio=3, zds=5, pattern="rslip", gyplussize=6, wyplussize=6),
    "gr1": dict(MICRO, gratio=1, pattern="rslip", gyplussize=4, wyplussize=4),
    "ragged": dict(xdv=13, ydv=7, zdv=9, gratio=2, zds=5, Rebs=20.0, ub=0.03, CFL=5.0,
                   pattern="pslip", gxplussize=2, wxplussize=24, gyplussize=5, wyplussize=2),
    # rows of 256 (coarse) and 512 (bands) nodes: the collide-stream's aligned-store form
    "wide_pslip": dict(xdv=256, ydv=8, zdv=10, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
                       pattern="pslip", gxplussize=8, wxplussize=8, gyplussize=4, wyplussize=4),
    "wide_sslip": dict(xdv=256, ydv=6, zdv=9, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=2.0,
                       pattern="sslip", gxplussize=3, wxplussize=5),
    # filtered injection (FLTR): the thinnest bands, whose filter layers reach the
    # wall layers, and deeper bands with posts
    "fltr": dict(MICRO, gratio=4, zds=6, pattern="rslip", gyplussize=4, wyplussize=4, fltr=1),
    "fltr_deep": dict(xdv=12, ydv=8, zdv=14, gratio=4, zds=8, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
                      pattern="pslip", gxplussize=8, wxplussize=8, gyplussize=4, wyplussize=4, fltr=1),
}
INIT = dict(init_kind="laminar_noise", seed=42, u_max=0.05, noise=0.05)
GRIDS = ("coarse", "lower", "upper")


def pair(kw, **extra):
    p = dnsg.make_params(**kw, **INIT, **extra)
    o = Oracle(p, threads=4)
    o.init_seeded()
    s = dnsg.Solver(p)
    st = o.state()
    s.upload(st["coarse"], st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
    return o, s


# '-i' start (initializer.c, inter_init.c): uniform grids finer than the fine grid with
# consistent extents, and coarser than it (one-sided z stencil at the lower wall)
UNIFORM = {
    "uniform_fine": (dict(MICRO, pattern="rslip", gyplussize=4, wyplussize=4), (64, 32, 50)),
    "uniform_coarse": (dict(MICRO), (24, 12, 13)),
}


@pytest.mark.parametrize("name", list(UNIFORM))
def test_uniform_start_bitwise(name):
    """the interpolated lattices, the wall layers and the three interface levels equal the
    oracle's (itself pinned to the reference's step-0 checkpoint) bit for bit, and stay
    so for 25 constant-force steps"""
    kw, (xu, yu, zu) = UNIFORM[name]
    uni = uniform_field(xu, yu, zu)
    p = dnsg.make_params(**kw, force_mode="const", g_const=2e-5)
    o = Oracle(p, threads=4)
    o.init_uniform(uni, xu, yu, zu)
    s = dnsg.Solver(p)
    print(f"uniform start {name}: {xu}x{yu}x{zu + 2}", flush=True)
    s.init_uniform(uni[0], xu, yu, zu)
    g, ref = s.download(), o.state()
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(g[k], ref[k]), f"{k}: max diff {np.max(np.abs(g[k] - ref[k]))}"
    assert abs(s.mass() - o.mass()) <= 1e-14
    dg = s.step(25)
    dr = o.run(25)
    g, ref = s.download(), o.state()
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(g[k], ref[k]), f"{k}: max diff {np.max(np.abs(g[k] - ref[k]))}"
    assert relmax(dg["flowrate"], dr["flowrate"]) <= 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_uniform_start_shifted_tiers(prec):
    """shifted storage of the same interpolated state: only the storage rounding differs"""
    kw, (xu, yu, zu) = UNIFORM["uniform_coarse"]
    uni = uniform_field(xu, yu, zu)
    o = Oracle(dnsg.make_params(**kw), threads=4)
    o.init_uniform(uni, xu, yu, zu)
    s = dnsg.Solver(dnsg.make_params(**kw, precision=prec))
    s.init_uniform(uni[0], xu, yu, zu)
    g, ref = s.download(), o.state()
    eps = {"fp32": 2e-8, "fp16": 2e-4}[prec]  # absolute, on f ~ w +- 1e-2
    for k in GRIDS:
        assert np.max(np.abs(g[k] - ref[k])) <= eps, k


def test_uniform_start_rejects_uncovered_grid():
    """a uniform grid whose interpolation stencils would leave the array (the reference
    reads out of bounds there) is refused"""
    kw, (xu, yu, zu) = UNIFORM["uniform_coarse"]
    uni = uniform_field(4, yu, zu)
    s = dnsg.Solver(dnsg.make_params(**kw))
    with pytest.raises(dnsg.DnsgError, match="does not cover"):
        s.init_uniform(uni[0], 4, yu, zu)
    o = Oracle(dnsg.make_params(**kw))
    with pytest.raises(ValueError, match="does not cover"):
        o.init_uniform(uni, 4, yu, zu)


@pytest.mark.parametrize("name", list(CASES))
def test_seeded_init_bitwise(name):
    p = dnsg.make_params(**CASES[name], **INIT)
    o = Oracle(p)
    o.init_seeded()
    s = dnsg.Solver(p)
    s.init_state()
    g, ref = s.download(), o.state()
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(g[k], ref[k]), k


def test_upload_download_roundtrip():
    o, s = pair(CASES["pslip"])
    g, ref = s.download(), o.state()
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(g[k], ref[k]), k


# band path: the handle's default (both gratio-2 substeps in one launch for FP64
# constant force), one launch per substep, one launch for both
BAND = ["default", "0", "1"]


@pytest.mark.parametrize("band", BAND)
@pytest.mark.parametrize("name", list(CASES))
def test_constant_force_bitwise(name, band, monkeypatch):
    if band != "default":
        monkeypatch.setenv("DNSG_BAND2", band)
    o, s = pair(CASES[name], force_mode="const", g_const=2e-5)
    n = 25
    dg = s.step(n)
    dr = o.run(n)
    g, ref = s.download(), o.state()
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(g[k], ref[k]), f"{k}: max diff {np.max(np.abs(g[k] - ref[k]))}"
    for f in ("flowrate", "density"):
        assert relmax(dg[f], dr[f]) <= 1e-12, f
    for f in ("str_up", "str_down", "stress"):
        assert relmax(dg[f], dr[f]) <= 1e-10, f
    assert np.array_equal(dg["pgrad"], dr["pgrad"])
    assert np.array_equal(dg["time"], dr["time"])


def test_constant_force_bitwise_split_directions():
    """a launch of 128 interface tiles (256 x 64 coarse columns, both bands): the FP64
    fill shares each tile's 19 directions over about 6 workgroups (the small cases
    above take one direction per workgroup, the full-size one all 19), bit for bit"""
    kw = dict(xdv=256, ydv=64, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
              pattern="rslip", gyplussize=8, wyplussize=8)
    o, s = pair(kw, force_mode="const", g_const=2e-5)
    n = 10
    s.step(n)
    o.run(n)
    g, ref = s.download(), o.state()
    for k in GRIDS + ("lin", "uin"):
        assert np.array_equal(g[k], ref[k]), f"{k}: max diff {np.max(np.abs(g[k] - ref[k]))}"


@pytest.mark.parametrize("band", BAND)
@pytest.mark.parametrize("name", list(CASES))
def test_flux_control_parity(name, band, monkeypatch):
    if band != "default":
        monkeypatch.setenv("DNSG_BAND2", band)
    o, s = pair(CASES[name])
    n = 50
    dg = s.step(n)
    dr = o.run(n)
    g, ref = s.download(), o.state()
    for k in GRIDS:
        assert relmax(g[k], ref[k]) <= 1e-12, k
    ub = s.params.ub
    assert np.max(np.abs(dg["force"] - dr["force"])) <= 1e-12 * ub
    for f in ("flowrate", "density", "momentum"):
        assert relmax(dg[f], dr[f]) <= 1e-12, f
    for f in ("str_up", "str_down", "stress"):
        assert relmax(dg[f], dr[f]) <= 1e-10, f


def test_cfg1_200_steps_tier_a():
    """64x32x32 coarse, factor-2 bands of 8 layers, flux control, 200 steps."""
    kw = dict(xdv=64, ydv=32, zdv=32, gratio=2, zds=8, Rebs=182.5, ub=0.0333333333333, CFL=4.0)
    o, s = pair(kw)
    dg = s.step(200)
    dr = o.run(200)
    g, ref = s.download(), o.state()
    for k in GRIDS:
        assert relmax(g[k], ref[k]) <= 1e-12, k
    assert relmax(dg["flowrate"], dr["flowrate"]) <= 1e-12


@pytest.mark.parametrize("name", ["rslip", "fltr"])
@pytest.mark.parametrize("mode", ["const", "flux"])
def test_granular_entry_points_equal_step(mode, name):
    """the one-to-one call sequence of main:1256-1375 gives the fused step's state:
    bit for bit under constant force; under the flux controller the fused step takes
    <rho u> from the previous step's writes, a sum in another order (1e-13)"""
    kw = CASES[name]
    extra = dict(force_mode="const", g_const=2e-5) if mode == "const" else {}
    _, a = pair(kw, **extra)
    _, b = pair(kw, **extra)
    a.step(3)
    GR = b.dims.GR
    for ss in range(3):
        b.encalc2(ss)
        b.wshearstr(ss)
        b.xcomputations()
        for it in range(1, GR):
            b.xfcomputations()
            b.ctfdtransfer(it)
            b.fexchange()
        b.xfcomputations()
        b.ctfdtransfer(GR)
        b.ftcdtransfer()
        b.exchange()
        b.fexchange()
    ga, gb = a.download(), b.download()
    for k in GRIDS + ("lin", "uin"):
        if mode == "const":
            assert np.array_equal(ga[k], gb[k]), k
        else:
            assert relmax(ga[k], gb[k]) <= 1e-13, (k, relmax(ga[k], gb[k]))


@pytest.mark.parametrize("name", ["rslip", "wide_pslip", "gr3", "fltr"])
def test_flux_controller_from_writes(name):
    """G = ub - <rho u> summed from the previous step's writes (dnsg_step) against the
    same sum taken over the lattices before every step (state reloaded between steps)"""
    n = 30
    _, a = pair(CASES[name])
    da = a.step(n)
    _, b = pair(CASES[name])
    db = {k: [] for k in da}
    for _ in range(n):
        st = b.download()  # reloading the state makes the next step sum the lattices
        b.upload(st["coarse"], st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
        d = b.step(1)
        for k in db:
            db[k].append(d[k][0])
    ub = dnsg.make_params(**CASES[name]).ub
    assert np.max(np.abs(da["force"] - np.array(db["force"]))) <= 1e-13 * ub
    for k in ("momentum", "density", "flowrate"):
        assert relmax(da[k], np.array(db[k])) <= 1e-13, k
    ga, gb = a.download(), b.download()
    for k in GRIDS:
        assert relmax(ga[k], gb[k]) <= 1e-12, k


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_velocity_snapshot_is_vel_print(prec):
    """dnsg_velocity_snapshot (formed on the device) against vel_print.c:86-117
    evaluated in its own operation order on the downloaded state, wall layers
    included: bit for bit"""
    kw = dict(CASES["rslip"], force_mode="const", g_const=2e-5, precision=prec)
    _, s = pair(kw)
    s.step(4)
    st = s.download()
    E = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, 1, -1, -1, 0, 0, 0, 0],
                  [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1],
                  [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]], dtype=float)
    for g, key in enumerate(GRIDS):
        f = st[key]
        u = [np.zeros(f.shape[:3]) for _ in range(3)]
        dens = np.zeros(f.shape[:3])
        for a in range(19):
            for c in range(3):
                u[c] = u[c] + E[c, a] * f[..., a]
            dens = dens + f[..., a]
        ok = dens > 0.0
        vel = np.stack([np.where(ok, uc / np.where(ok, dens, 1.0), 0.0) for uc in u], axis=-1)
        vg, dg = s.velocity(g)
        assert np.array_equal(vg, vel), key
        assert np.array_equal(dg, np.where(ok, dens, 0.0)), key


@pytest.mark.parametrize("name", ["pslip", "gr3"])
def test_mass_and_statistics(name):
    o, s = pair(CASES[name])
    s.step(10)
    o.run(10)
    assert abs(s.mass() - o.mass()) <= 1e-13 * abs(o.mass())
    sg, so = s.statistics(), o.statistics()
    for r in range(9):
        assert relmax(sg[r], so[r]) <= 1e-11, r


def test_deterministic_rerun():
    _, a = pair(CASES["pslip"])
    _, b = pair(CASES["pslip"])
    da, db = a.step(20), b.step(20)
    for f in dnsg.DIAG_FIELDS:
        assert np.array_equal(da[f], db[f]), f
    ga, gb = a.download(), b.download()
    for k in GRIDS:
        assert np.array_equal(ga[k], gb[k]), k


def test_nan_in_state_aborts_with_status():
    """a NaN planted in the near-wall band is reported as the reference's abort"""
    o, s = pair(CASES["none"])
    st = o.state()
    st["lower"][3, 2, 2, 7] = np.nan
    s.upload(st["coarse"], st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
    with pytest.raises(dnsg.DnsgError) as ei:
        s.step(1)
    assert ei.value.rc == -3


@pytest.mark.parametrize("timed", [False, True])
def test_nan_stops_at_the_failing_step(timed):
    """A NaN planted in the middle of the coarse grid reaches the near-wall band layers
    a few steps later.  The oracle stops at the top of the step whose wall stress is
    NaN, as the reference aborts there (wshearstr.c:139-143).  The GPU run, asked for
    many more steps in one call, returns the NaN status with the records up to that
    step equal to the oracle's, and holds the state of that step's top: every lattice
    and interface level equal to the oracle's wherever the oracle's is finite
    (constant force, bitwise), and a further call fails again at the same step.
    Where the oracle holds a NaN the GPU may hold a finite value: the interface fill
    drops the products of a coefficient with t = 0 or u = 0 (the fine columns on the
    low x and low y edges of each coarse cell), so a NaN coefficient does not reach
    those columns as it does in ctfdtransfer_opt.c's full Horner form.  No NaN
    appears on the GPU where the oracle has none."""
    import ctypes
    kw = dict(CASES["rslip"], force_mode="const", g_const=2e-5)
    o, s = pair(kw)
    st = o.state()
    st["coarse"][8, 4, 6, 0] = np.nan
    s.upload(st["coarse"], st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
    o2 = Oracle(dnsg.make_params(**kw, **INIT), threads=4)
    o2.init_seeded()
    for a, aid in (("coarse", 1), ("lower", 3), ("upper", 5)):
        o2.array(aid)[...] = st[a]
    o2.set_levels(0, st["lin"])
    o2.set_levels(1, st["uin"])
    recs, fail = [], None
    for ss in range(40):
        rc, d = o2.step(ss)
        recs.append(d)
        if rc:
            assert rc == -3
            fail = ss
            break
    assert fail is not None and fail >= 2, fail
    n = 60
    d = (dnsg.Diag * n)()
    if timed:
        ms = ctypes.c_double()
        rc = s.lib.dnsg_run_timed(s.h, n, d, ctypes.byref(ms))
    else:
        rc = s.lib.dnsg_step(s.h, n, d)
    print(f"NaN planted: oracle fails at step {fail}, GPU rc {rc}", flush=True)
    assert rc == -3
    assert f"step {fail}" in s.lib.dnsg_last_error(s.h).decode()
    for k in range(fail + 1):
        for f in ("time", "force", "density", "flowrate"):
            a, b = getattr(d[k], f), getattr(recs[k], f)
            assert (np.isnan(a) and np.isnan(b)) or a == b, (k, f, a, b)
    assert np.isnan(d[fail].stress) and not np.isnan(d[fail - 1].stress)
    g, ref = s.download(), o2.state()
    for k in GRIDS + ("lin", "uin"):
        fin = np.isfinite(ref[k])
        assert np.array_equal(g[k][fin], ref[k][fin]), k
        assert not np.isnan(g[k][~np.isnan(ref[k])]).any(), k
    assert np.isnan(g["coarse"]).any()
    d1 = (dnsg.Diag * 1)()
    assert s.lib.dnsg_step(s.h, 1, d1) == -3
    assert d1[0].time == d[fail].time


def test_running_statistics_match_host_average():
    """time averages = host average of the per-snapshot profiles (law of total variance)"""
    o, s = pair(CASES["rslip"])
    raw = []
    for _ in range(4):
        s.step(3)
        s.stats_accumulate()
        raw.append(s.statistics())
    avg, cnt = s.stats_average()
    assert cnt == 4
    R = np.array(raw)
    means = R[:, [0, 1, 2, 7]].mean(axis=0)
    assert relmax(avg[[0, 1, 2, 7]], means) <= 1e-12
    for row, mrow in ((3, 0), (4, 1), (5, 2)):
        var = (R[:, row] ** 2).mean(axis=0) + R[:, mrow].var(axis=0)
        assert relmax(avg[row] ** 2, var) <= 1e-10, row
    dvar = (R[:, 8] ** 2).mean(axis=0) + (R[:, 7] - 1.0).var(axis=0)
    assert relmax(avg[8] ** 2, dvar) <= 1e-10
    uw = R[:, 6].mean(axis=0) + (R[:, 0] * R[:, 2]).mean(axis=0) - R[:, 0].mean(axis=0) * R[:, 2].mean(axis=0)
    assert relmax(avg[6], uw) <= 1e-9
