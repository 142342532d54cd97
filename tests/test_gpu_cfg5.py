"""The Re_tau = 395 square-post channel in shifted FP32 (coarse 1024 x 512 x 256, factor-2
bands 2 x 2048 x 1024 x 33, solid posts 32 x 32 fine in a 64 x 64 period, 402.7M lattice
updates per coarse step, about 41 GB of FP32 lattices), bench.py's cfg5_params.

* full size: the shifted-FP32 handle against the FP64 plain handle of the same channel
  (82 GB, also resident on one MI355X; the plain path equals the oracle bit for bit at
  every size the parity tests cover) after 3 coarse steps from the same seeded start:
  every reported scalar finite, the FP64 mass of the FP32 state within 1e-6 of the
  plain path's (FP32 storage of g = f - w rounds each value at about 3e-9), and the
  mean streamwise profile within 5e-4 of its maximum (the precision tier's bar);
* the same tau, post pattern and band depth at a reduced periodic extent (64 x 32 coarse
  columns): the FP32 state node by node within 1e-7 of the FP64 oracle after 2 steps,
  the bar of the precision tests (FP32 arithmetic leaves about 1e-8 on f, a wrong
  interpolation weight or wall rule about 1e-4);
* the three grids' 1-D spectra at full size are finite and non-negative.
The node codes of the full-size channel are checked against the reference's own
solid_init / slip_posts_init dump in test_gpu_masks.py."""
import os
import sys

import numpy as np
import pytest

from helpers import Oracle, dnsg, relmax

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _run(prec, n):
    s = dnsg.Solver(bench.cfg5_params(dnsg, prec))
    s.init_state()
    d = s.step(n)
    out = dict(diag=d, mass=s.mass(), stats=s.statistics())
    return s, out


def test_full_size_fp32_against_fp64():
    print("cfg5 full size: fp64 plain, 3 steps", flush=True)
    s, ref = _run("fp64", 3)
    s.close()
    print("cfg5 full size: fp32 shifted, 3 steps", flush=True)
    s, got = _run("fp32", 3)
    for k in ("flowrate", "density", "stress", "str_up", "str_down"):
        assert np.isfinite(got["diag"][k]).all(), k
    assert np.isfinite(got["stats"]).all()
    print(f"cfg5 mass fp64 {ref['mass']:.15f} fp32 {got['mass']:.15f}", flush=True)
    assert abs(got["mass"] - ref["mass"]) <= 1e-6
    assert relmax(got["stats"][0], ref["stats"][0]) <= 5e-4
    assert relmax(got["diag"]["flowrate"], ref["diag"]["flowrate"]) <= 5e-4
    print("cfg5 full size: spectra of the three grids", flush=True)
    for g in range(3):
        ex, ey = s.spectra(g)
        assert np.isfinite(ex).all() and np.isfinite(ey).all(), g
        assert (ex >= 0).all() and (ey >= 0).all(), g
        assert ex.sum() > 0, g
    s.close()


def test_reduced_extent_fp32_against_oracle():
    p64 = bench.cfg5_params(dnsg, "fp64", xdv=64, ydv=32)
    o = Oracle(p64, threads=8)
    o.init_seeded()
    o.run(2)
    ref = o.state()
    s = dnsg.Solver(bench.cfg5_params(dnsg, "fp32", xdv=64, ydv=32))
    s.init_state()
    s.step(2)
    g = s.download()
    for k in ("coarse", "lower", "upper"):
        err = float(np.max(np.abs(g[k] - ref[k])))
        print(f"cfg5 64x32 columns {k}: max |f32 - f64| = {err:.2e}", flush=True)
        assert err <= 1e-7, (k, err)
