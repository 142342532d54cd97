"""Tier (b): the long run of the Re_tau = 180 stripe channel (bench.py's workload:
coarse 512 x 256 x 128, factor-2 bands, u_tau = 0.005, constant body force
G = u_tau^2 / h, laminar start at Dean's bulk velocity with seeded +-10% noise), 20,000
coarse steps in the FP64 plain path and in the shifted FP32 tier, with plane statistics
(statistics.c:48-292, on the device) every 200 steps after step 10,000.

Beyond a few hundred steps two runs of a chaotic flow decorrelate, so the tiers are
compared through their time averages: mean U+, the rms of u', v', w' and -u'w' (w
wall-normal), in wall units of the nominal u_tau = sqrt(G h).  The uncertainty of each
average is its batch-means standard error (5 batches of 10 samples); the bar is
    |FP32 - FP64| <= 3 sqrt(SE_32^2 + SE_64^2) + 1e-3 max|FP64 profile|
at every profile slot, the 1e-3 floor covering a flow whose batch means barely move.

Whatever the flow does, the streamwise momentum of the whole channel obeys
    dU_b/dt = G - tau_w / h
(body force against the mean wall shear of both walls, h the half-height in coarse
cells): the change of the bulk velocity over the averaging window must match the
time integral of G - tau_w/h from the per-step records within 2% of that change or
1e-3 of G times the window (the wall-stress stencil and the encalc2 volume weights
are second-order approximations).  Whether the flow became turbulent is printed
(the rms levels against the laminar start's decay)."""
import os
import sys

import numpy as np
import pytest

from helpers import dnsg

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

TOTAL, START, EVERY, NB = 20000, 10000, 200, 5


def long_run(prec):
    p = bench.workload_params(dnsg, prec)
    s = dnsg.Solver(p)
    s.init_state()
    d = s.dims
    h = (d.NZs - 2.0) / (2.0 * d.GR)
    G = p.g_const
    samples, rec, hist = [], {k: [] for k in ("momentum", "density", "stress")}, []
    done = 0
    while done < TOTAL:
        _, dg = s.run_timed(EVERY)
        for k in rec:
            rec[k].append(dg[k])
        done += EVERY
        st = s.statistics()
        if done % 2000 == 0:
            hist.append((done, float(st[3].max()), float(st[5].max())))
        if done > START:
            samples.append(st)
    s.close()
    for k in rec:
        rec[k] = np.concatenate(rec[k])
    return dict(samples=np.array(samples), rec=rec, hist=hist, h=h, G=G, utau=np.sqrt(G * h))


def profiles(run):
    """per-sample wall-unit profiles: U+, u'+, v'+ (wall-normal), w'+ (spanwise), -u'v'+"""
    S, ut = run["samples"], run["utau"]
    return dict(U=S[:, 0] / ut, u=S[:, 3] / ut, v=S[:, 5] / ut, w=S[:, 4] / ut, uv=-S[:, 6] / ut ** 2)


def batch_stats(x):
    b = x.reshape(NB, -1, x.shape[-1]).mean(axis=1)
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(NB)


@pytest.fixture(scope="module")
def runs():
    out = {}
    for prec in ("fp64", "fp32"):
        print(f"long run: {prec}, {TOTAL} steps, statistics every {EVERY} after {START}", flush=True)
        out[prec] = long_run(prec)
        print(f"long run: {prec} rms history (step, max u'rms, max w'rms) {out[prec]['hist']}", flush=True)
    return out


def test_fp32_within_statistical_uncertainty(runs):
    P64, P32 = profiles(runs["fp64"]), profiles(runs["fp32"])
    for k in ("U", "u", "v", "w", "uv"):
        assert np.isfinite(P32[k]).all() and np.isfinite(P64[k]).all(), k
        m64, se64 = batch_stats(P64[k])
        m32, se32 = batch_stats(P32[k])
        bar = 3.0 * np.sqrt(se32 ** 2 + se64 ** 2) + 1e-3 * np.max(np.abs(m64))
        diff = np.abs(m32 - m64)
        worst = int(np.argmax(diff / bar))
        print(f"long run {k}: max|fp64| {np.max(np.abs(m64)):.4g}, max|diff| {diff.max():.3g}, "
              f"max SE {max(se64.max(), se32.max()):.3g}, worst slot {worst} diff/bar {diff[worst] / bar[worst]:.3f}",
              flush=True)
        assert (diff <= bar).all(), (k, worst, diff[worst], bar[worst])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_bulk_momentum_balance(runs, prec):
    r = runs[prec]
    ub = r["rec"]["momentum"] / r["rec"]["density"]
    tau = r["rec"]["stress"]
    i0, i1 = START, TOTAL - 1   # records are at the top of steps 0 .. TOTAL-1
    change = ub[i1] - ub[i0]
    # trapezoid over the steps i0 .. i1 of the per-step rate G - tau_w / h
    rate = r["G"] - tau[i0:i1 + 1] / r["h"]
    predicted = float(np.sum(0.5 * (rate[:-1] + rate[1:])))
    print(f"long run {prec}: U_b {ub[i0]:.6f} -> {ub[i1]:.6f}, change {change:.4e}, "
          f"G - tau_w/h integrated {predicted:.4e}, G x window {r['G'] * (i1 - i0):.4e}", flush=True)
    assert abs(change - predicted) <= max(0.02 * abs(change), 1e-3 * r["G"] * (i1 - i0))
