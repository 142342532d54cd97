"""The precision tier on the 64x32x32 perturbed laminar channel (coarse 64 streamwise x
32 spanwise x 32 wall-normal, factor-2 bands over 4 coarse cells per wall, tau_c 0.52,
tau_f 0.54, constant body force 1e-6, rho = 1, parabolic profile u_max 0.05 with seeded
+-5% noise, seed 1234, 500 coarse steps): shifted FP32 and FP16-storage runs against the
FP64 plain path, whose 500-step state is first checked bit for bit against the oracle.

The quantities are bench.py's precision-tier report (the same function): mean U+ (each
run scaled by its own friction velocity), u', v' (wall-normal), w' rms and -u'v' over
the interior profile slots, the mean slip velocity at the walls and the bulk flux.
Errors are max |tier - FP64| over the profile relative to the FP64 profile's largest
magnitude.  Bars (measured values in brackets):
                 U+            rms            -u'v'          slip velocity  flux
  FP32 storage   5e-3 [4e-6]   2e-3 [1.8e-4]  3e-3 [3.5e-4]  1e-3 [2.6e-5]  1e-4 [1.9e-5]
  FP16 storage   5e-3 [2.7e-4] 0.1  [6.1e-2]  0.1  [5.2e-2]  1e-2 [1.2e-3]  5e-4 [5.9e-5]
U+ within 0.5% is the precision tier's target; the others bound the storage rounding
of the fluctuations, which FP16's 11-bit mantissa leaves at a few percent of the
decaying noise's rms."""
import os
import sys

import numpy as np
import pytest

from helpers import Oracle, dnsg, relmax

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

BARS = {"fp32": dict(U_plus=5e-3, rms=2e-3, uv=3e-3, slip=1e-3, flux=1e-4),
        "fp16": dict(U_plus=5e-3, rms=0.1, uv=0.1, slip=1e-2, flux=5e-4)}


def test_fp64_reference_run_equals_oracle():
    """the FP64 plain run the tiers are measured against is the oracle's, bit for bit"""
    p = bench.cfg_small_params(dnsg, "fp64")
    s = dnsg.Solver(p)
    s.init_state()
    dg = s.step(500)
    o = Oracle(p, threads=8)
    o.init_seeded()
    dr = o.run(500)
    g, ref = s.download(), o.state()
    for k in ("coarse", "lower", "upper", "lin", "uin"):
        assert np.array_equal(g[k], ref[k]), k
    assert relmax(dg["flowrate"], dr["flowrate"]) <= 1e-12


def test_precision_tier_bars():
    rep = bench.precision_tier(dnsg, steps=500)
    print(rep, flush=True)
    for prec, b in BARS.items():
        e = rep[prec]
        assert e["U_plus"]["rel_to_profile_max"] <= b["U_plus"], (prec, e["U_plus"])
        for k in ("u_rms", "v_rms", "w_rms"):
            assert e[k]["rel_to_profile_max"] <= b["rms"], (prec, k, e[k])
        assert e["uv"]["rel_to_profile_max"] <= b["uv"], (prec, e["uv"])
        assert e["slip_velocity"]["abs"] <= b["slip"] * abs(e["slip_velocity"]["fp64"]), (prec, e["slip_velocity"])
        assert e["flux_rel"] <= b["flux"], (prec, e["flux_rel"])
