"""Precision tier: distributions stored shifted (g = f - w*rho0) in FP32 or FP16,
collision and interface arithmetic in FP32, flux/mass/statistics in FP64.
Each shifted run starts from the same seeded state as an FP64 oracle run and its
profiles are compared with the oracle's after the run.

Tolerances (mean streamwise velocity profile, relative to its maximum):
  FP32 storage: 5e-4      FP16 storage: 2e-2
and the reported bulk flux within the same bounds.
"""
import numpy as np
import pytest

from helpers import Oracle, dnsg, relmax

pytestmark = pytest.mark.gpu

KW = dict(xdv=32, ydv=16, zdv=20, gratio=2, zds=6, tau_c=0.56, CFL=5.0, force_mode="const", g_const=2e-6,
          pattern="rslip", gyplussize=8, wyplussize=8, init_kind="laminar_noise", seed=11, u_max=0.05, noise=0.05)
TOL = {"fp32": 5e-4, "fp16": 2e-2}


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_shifted_tier_against_fp64_oracle(prec):
    n = 300
    o = Oracle(dnsg.make_params(**KW), threads=4)
    o.init_seeded()
    dr = o.run(n)
    s = dnsg.Solver(dnsg.make_params(**KW, precision=prec))
    s.init_state()
    dg = s.step(n)
    sg, so = s.statistics(), o.statistics()
    assert relmax(sg[0], so[0]) <= TOL[prec], relmax(sg[0], so[0])
    assert relmax(dg["flowrate"], dr["flowrate"]) <= TOL[prec]
    assert abs(s.mass() - o.mass()) <= (1e-6 if prec == "fp32" else 1e-4)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_shifted_tier_filtered_injection(prec):
    """the same bar with gratio 4 and the filtered injection (FLTR), whose thinnest
    bands filter over the wall layers"""
    kw = dict(KW, gratio=4, zds=6, fltr=1)
    n = 200
    o = Oracle(dnsg.make_params(**kw), threads=4)
    o.init_seeded()
    dr = o.run(n)
    s = dnsg.Solver(dnsg.make_params(**kw, precision=prec))
    s.init_state()
    dg = s.step(n)
    sg, so = s.statistics(), o.statistics()
    assert relmax(sg[0], so[0]) <= TOL[prec], relmax(sg[0], so[0])
    assert relmax(dg["flowrate"], dr["flowrate"]) <= TOL[prec]


@pytest.mark.parametrize("gratio", [1, 3])
def test_shifted_tier_other_gratios(gratio):
    """the FP32 bar at gratio 1 and 3: the shifted tiers' bicubic patch is a weight
    matrix formed at compile time per gratio (fine points at thirds for gratio 3)"""
    kw = dict(KW, gratio=gratio, zds=6)
    n = 200
    o = Oracle(dnsg.make_params(**kw), threads=4)
    o.init_seeded()
    dr = o.run(n)
    s = dnsg.Solver(dnsg.make_params(**kw, precision="fp32"))
    s.init_state()
    dg = s.step(n)
    sg, so = s.statistics(), o.statistics()
    assert relmax(sg[0], so[0]) <= TOL["fp32"], relmax(sg[0], so[0])
    assert relmax(dg["flowrate"], dr["flowrate"]) <= TOL["fp32"]


@pytest.mark.parametrize("gratio", [1, 2, 3, 4])
def test_fp32_state_after_two_steps(gratio):
    """node by node after two coarse steps from the seeded state: FP32 arithmetic
    leaves at most about 8e-9 on f (measured at gratio 1..4), while an interpolation
    weight at the wrong fine point moves the interface layer by about 1e-4 (5% noise
    on u = 0.05), so 1e-7 absolute separates the two; the band lattices include
    their interface layer"""
    kw = dict(KW, gratio=gratio, zds=6)
    o = Oracle(dnsg.make_params(**kw), threads=4)
    o.init_seeded()
    o.run(2)
    ref = o.state()
    s = dnsg.Solver(dnsg.make_params(**kw, precision="fp32"))
    s.init_state()
    s.step(2)
    g = s.download()
    for k in ("coarse", "lower", "upper"):
        err = float(np.max(np.abs(g[k] - ref[k])))
        print(f"gratio {gratio} {k}: max |f32 - f64| = {err:.2e}", flush=True)
        assert err <= 1e-7, (k, err)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_shifted_state_roundtrip(prec):
    """upload of an FP64 state into shifted storage and back loses only the storage rounding"""
    o = Oracle(dnsg.make_params(**KW))
    o.init_seeded()
    st = o.state()
    s = dnsg.Solver(dnsg.make_params(**KW, precision=prec))
    s.upload(st["coarse"], st["lower"], st["upper"], st["lin"], st["uin"], st["force"])
    g = s.download()
    eps = {"fp32": 2e-8, "fp16": 2e-4}[prec]  # absolute, on f ~ w +- 1e-2
    for k in ("coarse", "lower", "upper"):
        assert np.max(np.abs(g[k] - st[k])) <= eps, k
