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
