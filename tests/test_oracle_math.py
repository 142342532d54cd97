"""Closed-form checks of the oracle's functions, independent of the reference.

Each test is chosen so that a plausible slip in the restatement -- a dropped
term, a wrong sign or index, a transposed operand -- breaks it:
  * D3Q19 tables: isotropy of the weights, BI = 3w, opposite = -e, mirror = z-flip;
  * equilibrium: zeroth/first/second moments of feq;
  * collision: conservation of mass and momentum, momentum gain rho*G per step;
  * Hermite in time: exact for quadratics in time, t = 1 returns the newest level;
  * 4th-order derivatives: the exact Fourier symbol on periodic sinusoids;
  * bicubic coefficients: exact for bicubic polynomials given exact derivatives;
  * coarse->fine fill: f - feq equals tau_f/(GR tau_c) f_neq for uniform planes;
  * wall stress: the one-sided stencil is the exact wall slope of a quadratic;
  * composite physics: halfway bounce-back walls reach the Poiseuille profile,
    all-gas (specular) walls transmit no friction, the rest state has mass 1;
  * the counter-based PRNG: known-answer values.
"""
import ctypes

import numpy as np
import pytest

from helpers import ORC_LIN, ORC_LINT, ORC_L_FT, ORC_U_FT, Oracle, dnsg, oracle_lib, relmax

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(D)


@pytest.fixture(scope="module")
def tables(oracle_built):
    lib = oracle_lib()
    w = np.zeros(19); bi = np.zeros(19)
    e = np.zeros(57, dtype=np.int32); opp = np.zeros(19, dtype=np.int32); mir = np.zeros(19, dtype=np.int32)
    lib.orc_tables(_d(w), _d(bi), e.ctypes.data_as(I), opp.ctypes.data_as(I), mir.ctypes.data_as(I))
    return w, bi, e.reshape(3, 19), opp, mir


def test_lattice_tables(tables):
    w, bi, e, opp, mir = tables
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(e @ w, 0, atol=1e-16)
    second = np.einsum("q,aq,bq->ab", w, e, e)
    assert np.allclose(second, np.eye(3) / 3, atol=1e-16)
    fourth = np.einsum("q,aq,bq,cq,dq->abcd", w, e, e, e, e)
    delta = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", delta, delta) + np.einsum("ac,bd->abcd", delta, delta)
           + np.einsum("ad,bc->abcd", delta, delta)) / 9
    assert np.allclose(fourth, iso, atol=1e-16)
    assert np.array_equal(bi[1:], 3 * w[1:]) and bi[0] == 0
    assert np.array_equal(e[:, opp], -e)
    assert np.array_equal(e[:, mir], e * np.array([[1], [1], [-1]]))
    assert len({tuple(c) for c in e.T}) == 19


def test_equilibrium_moments(tables, oracle_built):
    w, _, e, _, _ = tables
    lib = oracle_lib()
    rng = np.random.default_rng(1)
    for _ in range(20):
        rho = 1 + 0.1 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(3)
        f = np.zeros(19)
        lib.orc_feq_node(rho, _d(np.ascontiguousarray(u)), _d(f))
        assert abs(f.sum() - rho) < 1e-14
        assert np.allclose(e @ f, rho * u, atol=1e-15)
        P = np.einsum("aq,bq,q->ab", e, e, f)
        assert np.allclose(P, rho * (np.eye(3) / 3 + np.outer(u, u)), atol=1e-14)


def test_collision_conserves_and_forces(tables, oracle_built):
    w, _, e, _, _ = tables
    lib = oracle_lib()
    rng = np.random.default_rng(2)
    for G in (0.0, 3e-4):
        for omega in (0.7, 1.3, 1.9):
            f = w * (1 + 0.05 * rng.standard_normal(19))
            rho, mom = f.sum(), e @ f
            g = f.copy()
            lib.orc_collide_node(_d(g), G, omega)
            assert abs(g.sum() - rho) < 1e-15
            assert abs((e @ g)[0] - (mom[0] + rho * G)) < 1e-15
            assert np.allclose((e @ g)[1:], mom[1:], atol=1e-16)
    # omega = 1 relaxes straight to equilibrium
    f = w * (1 + 0.05 * rng.standard_normal(19))
    u = (e @ f) / f.sum()
    feq = np.zeros(19)
    lib.orc_feq_node(f.sum(), _d(np.ascontiguousarray(u)), _d(feq))
    lib.orc_collide_node(_d(f), 0.0, 1.0)
    assert np.allclose(f, feq, atol=1e-16)


MICRO = dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4, Rebs=20.0, ub=0.0333333333333, CFL=5.0)


def test_hermite_exact_for_quadratics_in_time(oracle_built):
    for GR in (2, 3):
        o = Oracle(dnsg.make_params(**dict(MICRO, gratio=GR, zds=5)))
        rng = np.random.default_rng(3)
        shape = (o.dims.Xc, o.dims.Yc, 23)
        a, b, c = (rng.standard_normal(shape) for _ in range(3))
        q = lambda t: a + b * t + c * t * t  # levels at t = 1 (newest), 0, -1
        for upper in (0, 1):
            o.set_levels(upper, np.stack([q(1.0), q(0.0), q(-1.0)]))
        for it in range(1, GR + 1):
            o.lib.orc_tinterpolate_hrmt(o.o, it)
            lint = o.array(ORC_LINT)[0]
            assert relmax(lint, q(it / GR)) < 1e-14
        assert np.array_equal(o.array(ORC_LINT)[0], q(1.0))


def test_fourth_order_derivatives_fourier_symbol(oracle_built):
    o = Oracle(dnsg.make_params(**MICRO))
    X, Y = o.dims.Xc, o.dims.Yc
    lint = o.array(ORC_LINT)
    x = np.arange(X)[:, None, None]
    y = np.arange(Y)[None, :, None]
    kx, ky = 2 * np.pi * 3 / X, 2 * np.pi * 1 / Y
    sym = lambda t: (4 / 3) * np.sin(t) - (1 / 6) * np.sin(2 * t)
    lint[0] = np.sin(kx * x) * np.cos(ky * y) * np.ones((1, 1, 23))
    o.lib.orc_derivative(o.o, o.lib.orc_array(o.o, ORC_LINT))
    assert relmax(lint[1], sym(kx) * np.cos(kx * x) * np.cos(ky * y) * np.ones((1, 1, 23))) < 1e-13
    assert relmax(lint[2], -sym(ky) * np.sin(kx * x) * np.sin(ky * y) * np.ones((1, 1, 23))) < 1e-13
    assert relmax(lint[3], -sym(kx) * sym(ky) * np.cos(kx * x) * np.sin(ky * y) * np.ones((1, 1, 23))) < 1e-13


def test_bicubic_coefficients_exact(oracle_built):
    o = Oracle(dnsg.make_params(**MICRO))
    lint = o.array(ORC_LINT)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((23, 4, 4))  # p_v(x, y) = sum a_ij x^i y^j on the cell (ic, jc)
    ic, jc = 5, 3
    for (dx, dy) in ((0, 0), (1, 0), (1, 1), (0, 1)):
        for v in range(23):
            a = A[v]
            px = [[a[i, j] * (dx ** i) * (dy ** j) for j in range(4)] for i in range(4)]
            val = sum(px[i][j] for i in range(4) for j in range(4))
            ddx = sum(i * a[i, j] * dx ** max(i - 1, 0) * dy ** j for i in range(1, 4) for j in range(4))
            ddy = sum(j * a[i, j] * dx ** i * dy ** max(j - 1, 0) for i in range(4) for j in range(1, 4))
            dxy = sum(i * j * a[i, j] * dx ** max(i - 1, 0) * dy ** max(j - 1, 0)
                      for i in range(1, 4) for j in range(1, 4))
            lint[0, ic + dx, jc + dy, v] = val
            lint[1, ic + dx, jc + dy, v] = ddx
            lint[2, ic + dx, jc + dy, v] = ddy
            lint[3, ic + dx, jc + dy, v] = dxy
    c = np.zeros(23 * 16)
    o.lib.orc_bcucof(o.o, o.lib.orc_array(o.o, ORC_LINT), ic, jc, _d(c))
    assert relmax(c.reshape(23, 4, 4), A) < 1e-13


def test_coarse_to_fine_rescaling(tables, oracle_built):
    """uniform planes: the fine interface layer is feq + tau_f/(GR tau_c) f_neq"""
    w, _, e, _, _ = tables
    o = Oracle(dnsg.make_params(**MICRO))
    rho, u = 1.02, np.array([0.03, -0.01, 0.005])
    fneq = 1e-3 * np.random.default_rng(5).standard_normal(19)
    plane = np.concatenate([[rho], u, fneq])
    lv = np.broadcast_to(plane, (3, o.dims.Xc, o.dims.Yc, 23)).copy()
    for upper in (0, 1):
        o.set_levels(upper, lv)
    o.lib.orc_tinterpolate_hrmt(o.o, 1)
    o.lib.orc_derivative(o.o, o.lib.orc_array(o.o, ORC_LINT))
    o.lib.orc_derivative(o.o, o.lib.orc_array(o.o, 9))
    o.lib.orc_ctfdtransfer(o.o, 1)
    feq = np.zeros(19)
    o.lib.orc_feq_node(rho, _d(np.ascontiguousarray(u)), _d(feq))
    r = o.dims.omega_c / (o.dims.GR * o.dims.omega_f)  # = tau_f / (GR tau_c)
    assert abs(r - o.dims.tau_f / (o.dims.GR * o.dims.tau_c)) < 1e-15
    zdf = o.dims.zdf
    lower = o.array(ORC_L_FT)[:, :, zdf - 1, :]
    upper = o.array(ORC_U_FT)[:, :, 0, :]
    assert relmax(lower - feq, np.broadcast_to(r * fneq, lower.shape)) < 1e-11
    assert relmax(upper - feq, np.broadcast_to(r * fneq, upper.shape)) < 1e-11


def test_wall_stress_stencil_exact_for_quadratic(tables, oracle_built):
    """u_x(d) = a + b d + c d^2 near both walls -> stress = nu_f * rho * b"""
    o = Oracle(dnsg.make_params(**MICRO))
    a, b, c, rho = 0.001, 2e-3, -3e-5, 1.01
    zdf = o.dims.zdf
    lo, up = o.array(ORC_L_FT), o.array(ORC_U_FT)
    for k in range(zdf):
        for arr, d in ((lo, k - 0.5), (up, (zdf - 1) - k - 0.5)):
            f = np.zeros(19)
            o.lib.orc_feq_node(rho, _d(np.array([a + b * d + c * d * d, 0.0, 0.0])), _d(f))
            arr[:, :, k, :] = f
    dg = dnsg.Diag()
    o.lib.orc_wshearstr(o.o, 0, ctypes.byref(dg))
    nu = (1 / 6) * (2 / o.dims.omega_f - 1)
    assert abs(dg.str_down - nu * rho * b) < 1e-14
    assert abs(dg.str_up - nu * rho * b) < 1e-14


def test_rest_state_mass_is_one(oracle_built):
    o = Oracle(dnsg.make_params(**MICRO))
    o.init_rest()
    assert abs(o.mass() - 1.0) < 1e-14


def test_poiseuille_with_halfway_bounce_back(oracle_built):
    """no-slip walls + constant force -> u = G_f/(2 nu_f) d (2h - d) in fine units"""
    kw = dict(xdv=4, ydv=2, zdv=12, gratio=2, zds=4, tau_c=0.75, CFL=5.0, force_mode="const",
              g_const=1e-6, init_kind="rest")
    o = Oracle(dnsg.make_params(**kw))
    o.init_rest()
    o.run(4000)
    st = o.statistics()
    d, NZs = o.dims, o.dims.NZs
    nu_f = (d.tau_f - 0.5) / 3
    G_f = 1e-6 / d.GR
    H = NZs - 2  # 2h in fine units; walls halfway at 0 and H
    zdf, zdc, GR = d.zdf, d.zdc, d.GR
    dist = np.zeros(d.NZ)
    dist[1:zdf] = np.arange(1, zdf) - 0.5
    dist[zdf:zdf + zdc - 4] = GR * (np.arange(2, zdc - 2) - 1) + (zdf - 1.5)
    dist[zdf + zdc - 4:2 * zdf + zdc - 5] = np.arange(0, zdf - 1) + GR * (zdc - 3) + zdf - 1.5
    used = dist > 0
    exact = G_f / (2 * nu_f) * dist * (H - dist)
    # velocities in the statistics profile are in each grid's own lattice units:
    # coarse-slot velocities are coarse units = fine units (acoustic scaling)
    err = np.max(np.abs(st[0][used] - exact[used])) / exact.max()
    assert err < 0.02, err


def _plug_flow(o, U):
    """fill every lattice and the interface levels with feq(1, (U, 0, 0))"""
    f = np.zeros(19)
    o.lib.orc_feq_node(1.0, _d(np.array([U, 0.0, 0.0])), _d(f))
    for aid in range(6):
        o.array(aid)[...] = f
    plane = np.concatenate([[1.0, U, 0.0, 0.0], np.zeros(19)])
    for upper in (0, 1):
        o.set_levels(upper, np.broadcast_to(plane, (3, o.dims.Xc, o.dims.Yc, 23)).copy())
    return f


def test_constant_force_mode_is_the_controller_path(oracle_built):
    """The constant-force mode (the bitwise GPU cases) and the reference's flux
    controller differ only in where the body force comes from: encalc2.c:129 forms
    G = ub - <rho u>, the constant mode takes g_const.  The controller's run is pinned
    byte for byte to the reference's own files (test_oracle_golden.py).  Given, step by
    step, the G the controller formed, the constant-force oracle reproduces that run bit
    for bit, walls, interface and injection included; a G one ulp off at one step does
    not."""
    kw = dict(MICRO, pattern="rslip", gyplussize=4, wyplussize=4, init_kind="laminar_noise", seed=7,
              u_max=0.05, noise=0.05)
    flux = Oracle(dnsg.make_params(**kw))
    const = Oracle(dnsg.make_params(**kw, force_mode="const", g_const=3e-5))
    bent = Oracle(dnsg.make_params(**kw, force_mode="const", g_const=3e-5))
    for o in (flux, const, bent):
        o.init_seeded()
    gs = []
    for s in range(12):
        rc, d = flux.step(s)
        assert rc == 0
        gs.append(d.force)
        assert const.step_forced(s, d.force)[0] == 0
        assert bent.step_forced(s, np.nextafter(d.force, np.inf) if s == 5 else d.force)[0] == 0
    assert len(set(gs)) > 1  # the controller moved the force
    a, b, c = flux.state(), const.state(), bent.state()
    for k in ("coarse", "lower", "upper", "lin", "uin"):
        assert np.array_equal(a[k], b[k]), k
    assert any(not np.array_equal(a[k], c[k]) for k in ("coarse", "lower", "upper"))


def test_specular_walls_keep_plug_flow_exact(oracle_built):
    """all-gas walls transmit no shear: uniform plug flow is an exact steady state,
    while the same flow between no-slip (bounce-back) walls is decelerated"""
    kw = dict(xdv=8, ydv=4, zdv=12, gratio=2, zds=4, tau_c=0.8, CFL=5.0, force_mode="const", g_const=0.0)
    gas = Oracle(dnsg.make_params(**kw, pattern="rslip", gyplussize=8, wyplussize=0))
    f0 = _plug_flow(gas, 0.04)
    gas.run(20)
    for aid in (1, 3, 5):  # ftemp of coarse, lower, upper
        arr = gas.array(aid)
        if aid == 3:
            arr = arr[:, :, 1:]
        if aid == 5:
            arr = arr[:, :, :-1]
        assert relmax(arr, np.broadcast_to(f0, arr.shape)) < 1e-14
    solid = Oracle(dnsg.make_params(**kw))
    _plug_flow(solid, 0.04)
    solid.run(20)
    st = solid.statistics()
    assert st[0][1] < 0.9 * 0.04  # the first fluid layer next to a no-slip wall slows down


def test_prng_known_answers(oracle_built):
    lib = oracle_lib()
    got = [lib.orc_prng_bits(42, g, c, n) for g, c, n in ((0, 0, 0), (0, 1, 0), (1, 0, 7), (2, 2, 123456789))]
    assert got == KAT_BITS
    u = np.array([lib.orc_prng_sym(9, 0, 0, n) for n in range(20000)])
    assert -1 <= u.min() and u.max() < 1 and abs(u.mean()) < 0.02 and abs(u.var() - 1 / 3) < 0.01


# splitmix64 finaliser (Steele, Lea, Flood 2014) composed as mix(mix(seed) ^ mix(grid<<56 ^ comp<<48 ^ node)),
# evaluated independently in Python
KAT_BITS = [14769503741126384973, 14157109251876740642, 11839393956308078277, 5352266596598019260]
