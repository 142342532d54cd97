#!/usr/bin/env python3
"""Benchmark of the grid-embedded D3Q19 channel time step on one MI355X.

Workload (the Re_tau = 180 stripe channel): coarse 512 x 256 x 128 (x
streamwise, y spanwise, z wall-normal, zdv = 128 including the two overlap
layers), factor-2 wall bands of 16 fluid layers (fine 1024 x 512 x 17 each),
streamwise superhydrophobic ridges on both walls (gas fraction 0.5, period 64
fine = 32 coarse cells), u_tau = 0.005, nu = u_tau h / 180 with h the true
half-height in coarse units ((NZs-2)/(2 GR) = 70.25), tau_c = 3 nu + 1/2,
tau_f = 2 (tau_c - 1/2) + 1/2, constant body force u_tau^2 / h, laminar
profile at the matching bulk velocity plus seeded +-10% noise.

A step is one coarse time step: coarse collide-stream, GR = 2 band substeps
with the interface transfers, the fine->coarse injection, and the per-step
flux / density / wall-stress reductions the reference logs every step.
LU per step = Nc + GR (N_lower + N_upper) = 50,331,648.

value = FP64 plain-path GLUPS with the lattices resident in HBM.  The line also
carries the shifted FP32 and FP16-storage tiers, the roofline of the dominant
kernel (band collide-stream, measured live with HIP events on the solver's
stream), the oracle (C restatement, OpenMP) timed on a bounded sample on
the host cores, and the slip-length serving queue: seeded requests arriving as
a Poisson stream, continuously batched, with p50/p95/p99 latency, throughput
and an FP64 re-solve of a subset.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(REPO, "dns-solvers-with-grid-embedding_amd")


def load_pkg():
    """the product binding (ctypes over libdnsg.so), loaded by path: the package
    directory's name is not a Python identifier"""
    import importlib.util
    spec = importlib.util.spec_from_file_location("dnsg_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod

METRIC = "GLUPS + %HBM roofline, FP32 vs FP64 at Re_tau=180; p50/p99 slip-queue latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_LU = {"fp64": 304, "fp32": 152, "fp16": 76}


def workload_params(mod, prec="fp64", force_mode="const"):
    GR, zds, zdv = 2, 16, 128
    zdf = zds + 1
    NZs = 2 * zdf + GR * (zdv - 4) + GR - 1
    h = (NZs - 2.0) / (2.0 * GR)       # coarse units
    u_tau = 0.005
    nu = u_tau * h / 180.0
    tau_c = 3.0 * nu + 0.5
    G = u_tau * u_tau / h
    ub = 15.6 * u_tau                   # bulk velocity of the laminar start (Dean: U_b ~ 15.6 u_tau)
    return mod.make_params(xdv=512, ydv=256, zdv=zdv, gratio=GR, zds=zds, tau_c=tau_c, ub=ub,
                           pattern="rslip", gyplussize=32, wyplussize=32, force_mode=force_mode,
                           g_const=G, precision=prec, init_kind="laminar_noise", seed=2024,
                           u_max=1.5 * ub, noise=0.10, CFL=0.1)


def cfg5_params(mod, prec="fp32", xdv=1024, ydv=512):
    """The Re_tau = 395 channel with square superhydrophobic posts: coarse 1024 (streamwise)
    x 512 (spanwise) x 256 (wall-normal, zdv incl. the two overlap layers), factor-2 bands
    of 32 fluid layers (fine 2048 x 1024 x 33 each), solid posts of 32 x 32 fine (16 x 16
    coarse) cells in a 64 x 64 fine period on both walls (solid fraction 0.25), shifted
    FP32 storage, the Re_tau = 180 channel's tau / forcing rule (u_tau = 0.005,
    nu = u_tau h / 395, tau_c = 3 nu + 1/2, G = u_tau^2 / h) and its seeded +-10% noise on
    the laminar profile at Dean's bulk velocity (U_b = 17.4 u_tau at Re_tau = 395).
    402,653,184 LU per coarse step; about 41 GB of FP32 lattices.  xdv, ydv shrink the
    periodic extents only (multiples of 32 keep the post pattern whole)."""
    GR, zds, zdv = 2, 32, 256
    zdf = zds + 1
    NZs = 2 * zdf + GR * (zdv - 4) + GR - 1
    h = (NZs - 2.0) / (2.0 * GR)
    u_tau = 0.005
    nu = u_tau * h / 395.0
    tau_c = 3.0 * nu + 0.5
    G = u_tau * u_tau / h
    ub = 17.4 * u_tau
    return mod.make_params(xdv=xdv, ydv=ydv, zdv=zdv, gratio=GR, zds=zds, tau_c=tau_c, ub=ub,
                           pattern="pslip", gxplussize=32, wxplussize=32, gyplussize=32, wyplussize=32,
                           force_mode="const", g_const=G, precision=prec, init_kind="laminar_noise",
                           seed=395, u_max=1.5 * ub, noise=0.10, CFL=0.1)


def run_cfg5(mod, steps, warmup):
    """cfg-5 leg: GLUPS of the full Re_tau = 395 posts channel in shifted FP32, the roofline
    of its band launch (HIP events on the solver's stream inside the timed steps), and the
    cost of the three grids' 1-D spectra, which the run takes every 500 steps."""
    p = cfg5_params(mod)
    s = mod.Solver(p)
    s.init_state()
    if warmup:
        s.run_timed(warmup)
    t0 = time.perf_counter()
    ms_ev, diag, sec = s.run_timed(steps, sections="band")
    wall = time.perf_counter() - t0
    d = s.dims
    lu = lu_per_step(d)
    band_lu = s.band_launch_substeps() * 2 * d.Xf * d.Yf * (d.zdf - 1)
    band_bytes = band_lu * BYTES_PER_LU["fp32"]
    gbs = band_bytes / (sec["band"] * 1e-3) / 1e9 if sec["band"] > 0 else None
    t1 = time.perf_counter()
    for g in range(3):
        ex, ey = s.spectra(g)
        if not (np.isfinite(ex).all() and np.isfinite(ey).all()):
            raise RuntimeError("cfg 5: non-finite spectra")
    spec_s = time.perf_counter() - t1
    ms_step = wall * 1e3 / steps
    out = dict(workload="Re_tau=395 square-post channel, coarse 1024x512x256, factor-2 bands 2x2048x1024x33, "
                        "posts 32x32 fine in a 64x64 period (solid fraction 0.25), shifted FP32, constant force",
               lu_per_step=lu, glups=round(lu * steps / wall / 1e9, 4), ms_per_step=round(ms_step, 4),
               steps=steps, warmup=warmup, hbm_roofline_frac_step=round(lu * steps / wall / 1e9 * 152 / HBM_PEAK_GBS, 4),
               band_launch=dict(lu=band_lu, bytes=band_bytes, ms=round(sec["band"], 5),
                                gbs=round(gbs, 1) if gbs else None,
                                frac=round(gbs / HBM_PEAK_GBS, 4) if gbs else None),
               spectra_s_three_grids=round(spec_s, 4),
               spectra_share_at_every_500_steps=round(spec_s / (500 * ms_step * 1e-3), 5),
               flowrate_last=float(diag["flowrate"][-1]), mass=s.mass())
    s.close()
    return out


def lu_per_step(d):
    return d.Xc * d.Yc * d.zdc + d.GR * 2 * d.Xf * d.Yf * (d.zdf - 1)


def run_gpu(mod, prec, steps, warmup, force_mode="const"):
    p = workload_params(mod, prec, force_mode)
    s = mod.Solver(p)
    s.init_state()
    if warmup:
        s.run_timed(warmup)
    # the timed steps carry HIP events just before and after each band launch (on the
    # solver's stream): the roofline's launch time comes from these same steps
    t0 = time.perf_counter()
    ms_ev, diag, sec = s.run_timed(steps, sections="band")   # synchronises the stream
    wall = time.perf_counter() - t0
    kt = s.kernel_times(3)   # every kernel group, back to back, in 3 further steps
    d = s.dims
    lu = lu_per_step(d)
    out = dict(glups=lu * steps / wall / 1e9, ms_per_step=wall * 1e3 / steps, ms_per_step_events=ms_ev / steps,
               kernels_ms=kt, lu_per_step=lu, flowrate_last=float(diag["flowrate"][-1]),
               stress_last=float(diag["stress"][-1]))
    subs = s.band_launch_substeps()                # 2: both gratio-2 substeps in one launch
    band_lu = subs * 2 * d.Xf * d.Yf * (d.zdf - 1)  # both bands, every substep the launch covers
    band_bytes = band_lu * BYTES_PER_LU[prec]
    band_ms = sec["band"]
    out["band_launch"] = dict(lu=band_lu, bytes=band_bytes, ms=band_ms, substeps=subs,
                              launches=steps * (1 if subs == 2 else d.GR - 1),
                              gbs=band_bytes / (band_ms * 1e-3) / 1e9 if band_ms > 0 else None)
    s.close()
    return out


TRAFFIC_PROFILES = {2: "pmc_band2_retau180_fp64.md",              # k_band2: both substeps in one launch
                    1: "pmc_collide_stream_retau180_fp64.md"}    # one collide-stream launch per substep


def traffic_from_profiles(substeps):
    """HBM bytes per band launch from the committed rocprofv3 --pmc summary of that
    band kernel, and a note.  The summary lists the kernel's sources and their
    SHA-256; a source that changed since makes the figure stale, and it is not used."""
    path = os.path.join(REPO, "profiles", TRAFFIC_PROFILES.get(substeps, ""))
    if not os.path.isfile(path):
        return None, "no counter summary for this band kernel"
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#") and not line.startswith("|"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    srcs = vals.get("sources", "").split()
    if not srcs or "sources_sha256" not in vals:
        return None, f"{os.path.basename(path)} names no source hash"
    h = hashlib.sha256()
    for f in srcs:
        fp = os.path.join(PKG_DIR, f)
        if not os.path.isfile(fp):
            return None, f"{f} (listed in {os.path.basename(path)}) is missing"
        h.update(open(fp, "rb").read())
    if h.hexdigest() != vals["sources_sha256"]:
        return None, f"stale: the sources listed in {os.path.basename(path)} changed since it was measured"
    try:
        return float(vals["hbm_bytes_per_launch"]), f"rocprofv3 FETCH_SIZE + WRITE_SIZE, {os.path.basename(path)}"
    except (KeyError, ValueError):
        return None, f"{os.path.basename(path)} has no hbm_bytes_per_launch"


def cpu_baseline(mod, nsteps):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import Oracle  # the oracle: C restatement of the reference step (the checker)
    from helpers import dnsg as omod  # its parameter structs come from the checker's own binding
    # threads: every core this process may run on, capped by the host's CPU share for one
    # GPU (OMP_NUM_THREADS where the host sets it); nproc and the affinity are reported
    nproc = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    share = int(os.environ.get("OMP_NUM_THREADS", str(avail)) or avail)
    cores = max(1, min(avail, share))
    p = workload_params(omod, "fp64")
    o = Oracle(p, threads=cores)
    o.init_seeded()
    o.step(0)  # first touch of every page outside the timed sample
    t0 = time.perf_counter()
    for s in range(1, nsteps + 1):
        o.step(s)
    dt = time.perf_counter() - t0
    lu = lu_per_step(o.dims)
    o.close()
    return dict(value=lu * nsteps / dt / 1e9, unit="GLUPS", cores=cores, kind="port", nproc=nproc,
                affinity_cores=avail, omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                sample=f"{nsteps} coarse steps of the same 512x256x128 + 2x1024x512x17 channel "
                       f"(oracle, C restatement of the reference step, OpenMP on {cores} threads, {dt:.1f} s)")


def cfg_small_params(mod, prec):
    """The 64x32x32 perturbed laminar channel: coarse 64 (streamwise) x 32 (spanwise) x 32
    (wall-normal), factor-2 bands over 4 coarse cells at each wall, tau_c 0.52 (tau_f 0.54),
    constant body force 1e-6, rho = 1, parabolic profile u_max 0.05 with seeded +-5% noise."""
    return mod.make_params(xdv=64, ydv=32, zdv=32, gratio=2, zds=8, tau_c=0.52, ub=0.05 * 2 / 3,
                           force_mode="const", g_const=1e-6, precision=prec, init_kind="laminar_noise",
                           seed=1234, u_max=0.05, noise=0.05, CFL=1.0)


def precision_tier(mod, steps=500):
    """Shifted FP32 and FP16-storage runs of the 64x32x32 channel against the FP64 plain
    path (which equals the oracle bit for bit under constant force): mean U+ (each run
    scaled by its own friction velocity from its wall stress), u', v', w' rms, -u'v'
    (v wall-normal) over the interior profile slots, the mean slip velocity at the walls
    (U extrapolated to the halfway wall from the first two fine layers, both walls
    averaged) and the bulk flux.  Errors are max |tier - fp64| over the profile, also
    relative to the FP64 profile's largest magnitude."""
    res, prof = {}, {}
    for prec in ("fp64", "fp32", "fp16"):
        s = mod.Solver(cfg_small_params(mod, prec))
        s.init_state()
        d = s.step(steps)
        st = s.statistics()
        NZ = s.dims.NZ
        s.close()
        st = np.asarray(st).reshape(9, NZ)
        utau = float(np.sqrt(abs(d["stress"][-1])))
        inner = slice(1, NZ - 1)
        U = st[0, inner] / utau
        us = 0.5 * ((1.5 * st[0, 1] - 0.5 * st[0, 2]) + (1.5 * st[0, NZ - 2] - 0.5 * st[0, NZ - 3]))
        prof[prec] = dict(U_plus=U, u_rms=st[3, inner] / utau, v_rms=st[5, inner] / utau, w_rms=st[4, inner] / utau,
                          uv=-st[6, inner] / utau ** 2, slip=us, flux=float(d["flowrate"][-1]), utau=utau)
    ref = prof["fp64"]
    for prec in ("fp32", "fp16"):
        t = prof[prec]
        e = {}
        for k in ("U_plus", "u_rms", "v_rms", "w_rms", "uv"):
            a = float(np.max(np.abs(t[k] - ref[k])))
            e[k] = dict(max_abs=a, rel_to_profile_max=a / max(float(np.max(np.abs(ref[k]))), 1e-300))
        e["slip_velocity"] = dict(value=t["slip"], fp64=ref["slip"], abs=abs(t["slip"] - ref["slip"]))
        e["flux_rel"] = abs(t["flux"] - ref["flux"]) / abs(ref["flux"])
        e["utau_rel"] = abs(t["utau"] - ref["utau"]) / ref["utau"]
        res[prec] = e
    res["workload"] = ("64x32x32 coarse, GR 2 bands 128x64x8, tau_c 0.52, g 1e-6, seeded laminar + 5% noise, "
                       f"{steps} steps; v'rms and -u'v' use the wall-normal component")
    res["fp64_u_tau"] = ref["utau"]
    res["fp64_slip_velocity"] = ref["slip"]
    return res


# The serving workload's envelope (the slip_workload defaults): streamwise and spanwise
# 32..256, wall-normal 16..64 coarse cells, Re_b 10..200.  "small" is a reduced envelope
# (wall-normal 16..32, Re_b 10..50) whose solves converge in seconds.
QUEUE_ENVELOPES = {"full": dict(x_range=(32, 256), y_range=(32, 256), z_range=(16, 64), re_range=(10.0, 200.0)),
                   "small": dict(x_range=(32, 128), y_range=(32, 128), z_range=(16, 32), re_range=(10.0, 50.0))}
QUEUE_STREAM = 256  # requests in the seeded stream; a run takes its first n


def run_queue(mod, n, rate, seed, n_check, envelope="full", max_steps=400000, first=0):
    """Slip-length serving queue: requests first .. first+n-1 of the seeded 256-request stream
    (mod.slip_workload over the envelope: bands on a random half, stripes or posts, gas
    fraction 0.2..0.9, period 8..64 fine cells, shifted FP32), Poisson arrivals at `rate`
    per second, each solved until its bulk velocity moves by < 1e-9 (relative) over
    1000 steps (or, at the FP32 storage floor, stays within 3e-5 over 5 windows).
    Latency = completion - arrival.  The first n_check of them are solved again in
    FP64 (plain path, the 1e-9 criterion) and their slip lengths compared."""
    import random
    env = QUEUE_ENVELOPES[envelope]
    work = mod.slip_workload(QUEUE_STREAM, seed, precision="fp32", **env)[first:first + n]
    n = len(work)
    rng = random.Random(seed + 1)
    t, arrivals = 0.0, []
    for _ in range(n):
        t += rng.expovariate(rate)
        arrivals.append(t)
    lu = []  # lattice updates per step of each request
    for d, p in work:
        dm = mod.dims_of(p)
        lu.append(dm.Xc * dm.Yc * dm.zdc + dm.GR * 2 * dm.Xf * dm.Yf * (dm.zdf - 1))
    q = mod.Queue(max_active=n, check_every=100, window=1000, tol=1e-9, max_steps=max_steps)
    t0 = q.clock()
    done, i, last = {}, 0, time.time()
    while len(done) < n:
        now = q.clock() - t0
        while i < n and arrivals[i] <= now:
            q.submit(work[i][1], i, t0 + arrivals[i])
            i += 1
        info = q.info()
        if info["active"] == 0 and info["pending"] == 0 and i < n:
            time.sleep(max(0.0, arrivals[i] - (q.clock() - t0)))
            continue
        for r in q.advance(100):
            done[r["id"]] = r
        if time.time() - last > 15:
            print(f"bench: queue t={q.clock() - t0:.0f}s submitted={i} done={len(done)} {info}", file=sys.stderr,
                  flush=True)
            last = time.time()
    wall = q.clock() - t0
    graphs = q.graph_launches()
    q.close()
    lat = np.array([done[k]["t_done"] - done[k]["t_arrival"] for k in range(n)])
    steps = np.array([done[k]["steps"] for k in range(n)])
    glu = sum(done[k]["steps"] * lu[k] for k in range(n)) / 1e9
    out = dict(requests=n, stream=f"requests {first}..{first + n - 1} of the {QUEUE_STREAM}-request stream of seed {seed}",
               arrival_rate_per_s=rate, seed=seed, precision="fp32 shifted, fp64 reductions",
               envelope=dict(name=envelope, **env), converged_tol=int(sum(done[k]["status"] == 0 for k in range(n))),
               steady_at_fp32_floor=int(sum(done[k]["status"] == 2 for k in range(n))),
               capped=int(sum(done[k]["status"] == 1 for k in range(n))),
               criterion="|dU/U| <= 1e-9 over 1000 steps, or (FP32) within 3e-5 of U 1..5 windows back",
               p50_s=float(np.percentile(lat, 50)), p95_s=float(np.percentile(lat, 95)),
               p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()), wall_s=wall,
               throughput_req_per_s=n / wall, glups=glu / wall, steps_median=int(np.median(steps)),
               steps_max=int(steps.max()))
    out["graph_intervals"] = graphs
    if n_check:
        q = mod.Queue(max_active=n_check, check_every=100, window=1000, tol=1e-9, max_steps=max_steps)
        for k in range(n_check):
            d = work[k][0]
            q.submit(mod.slip_request(d["xdv"], d["ydv"], d["zdv"], d["gratio"], d["pattern"], d["period"],
                                      d["gas_fraction"], d["re_b"], precision="fp64"), k)
        ref, last = {}, time.time()
        while len(ref) < n_check:
            for r in q.advance(1000):
                ref[r["id"]] = r
            if time.time() - last > 15:
                print(f"bench: queue fp64 check done={len(ref)}/{n_check} {q.info()}", file=sys.stderr, flush=True)
                last = time.time()
        q.close()
        db = [abs(done[k]["slip_length"] - ref[k]["slip_length"]) for k in range(n_check)]
        out["fp64_check"] = dict(requests=n_check, max_abs_diff_slip=float(max(db)),
                                 max_rel_diff_slip=float(max(db[k] / max(abs(ref[k]["slip_length"]), 1e-12)
                                                             for k in range(n_check))),
                                 fp64_converged_tol=int(sum(ref[k]["status"] == 0 for k in range(n_check))),
                                 per_request=[dict(id=first + k, slip_fp32=done[k]["slip_length"],
                                                   slip_fp64=ref[k]["slip_length"], status_fp32=done[k]["status"],
                                                   status_fp64=ref[k]["status"], steps_fp32=done[k]["steps"],
                                                   steps_fp64=ref[k]["steps"],
                                                   latency_s=done[k]["t_done"] - done[k]["t_arrival"],
                                                   lu_per_step=lu[k], **{f: work[k][0][f] for f in (
                                                       "xdv", "ydv", "zdv", "gratio", "pattern", "period",
                                                       "gas_fraction", "re_b")})
                                              for k in range(n_check)])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tiers", default="fp64,fp32,fp16")
    ap.add_argument("--queue-requests", type=int, default=4,
                    help="the first n requests of the seeded 256-request stream")
    ap.add_argument("--queue-rate", type=float, default=1.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--queue-check", type=int, default=4, help="requests re-solved in FP64 for the slip check")
    ap.add_argument("--queue-envelope", choices=sorted(QUEUE_ENVELOPES), default="full")
    ap.add_argument("--queue-first", type=int, default=0, help="first request of the stream to serve")
    ap.add_argument("--queue-only", action="store_true", help="run the serving queue only and print its record")
    ap.add_argument("--no-queue", action="store_true")
    ap.add_argument("--no-precision-report", action="store_true")
    ap.add_argument("--no-cfg5", action="store_true")
    ap.add_argument("--cfg5-steps", type=int, default=10)
    ap.add_argument("--pmc-run", action="store_true",
                    help="minimal run for hardware-counter passes: 3 FP64 steps of the workload, then one "
                         "1 GiB copy-probe launch whose known bytes calibrate FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the lattices of one channel live on a single "
              "MI355X); multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    mod = load_pkg()

    if args.pmc_run:
        s = mod.Solver(workload_params(mod, "fp64"))
        s.init_state()
        s.step(3)
        s.close()
        print(f"bench: copy probe {mod.hbm_probe(1 << 30, 0, 1):.1f} GB/s", file=sys.stderr, flush=True)
        return

    if args.queue_only:  # the serving workload alone: its record as one JSON line
        queue = run_queue(mod, args.queue_requests, args.queue_rate, 2025, min(args.queue_check, args.queue_requests),
                          args.queue_envelope, first=args.queue_first)
        print(json.dumps({"slip_queue": queue}), flush=True)
        return

    tiers = [t for t in args.tiers.split(",") if t]
    res = {}
    for prec in tiers:
        print(f"bench: {prec} {args.steps} steps (+{args.warmup} warmup)", file=sys.stderr, flush=True)
        res[prec] = run_gpu(mod, prec, args.steps, args.warmup)
        print(f"bench: {prec} {res[prec]['glups']:.3f} GLUPS, {res[prec]['ms_per_step']:.3f} ms/step, "
              f"kernels {res[prec]['kernels_ms']}", file=sys.stderr, flush=True)
    flux = None
    if "fp64" in tiers:  # the reference's own driving: G = ub - <rho u> every step
        print(f"bench: fp64 flux controller {args.steps} steps", file=sys.stderr, flush=True)
        r = run_gpu(mod, "fp64", args.steps, args.warmup, force_mode="flux")
        flux = {"glups": round(r["glups"], 4), "ms_per_step": round(r["ms_per_step"], 4),
                "kernels_ms": {a: round(b, 4) for a, b in r["kernels_ms"].items()}}
        print(f"bench: fp64 flux {r['glups']:.3f} GLUPS", file=sys.stderr, flush=True)
    probe = {k: round(mod.hbm_probe(4 << 30, i, 5), 1) for i, k in enumerate(("copy", "read", "write", "lattice"))}
    print(f"bench: HBM probe {probe} GB/s", file=sys.stderr, flush=True)
    main_t = res["fp64"] if "fp64" in res else res[tiers[0]]
    main_prec = "fp64" if "fp64" in res else tiers[0]
    bl = main_t["band_launch"]
    traffic, traffic_note = traffic_from_profiles(bl.get("substeps"))
    roof = dict(bound="hbm", achieved=bl["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=(bl["gbs"] / HBM_PEAK_GBS) if bl["gbs"] else None, traffic=traffic, traffic_source=traffic_note,
                kernel=("k_band2<double, false> (both wall bands, both gratio-2 substeps in one launch)"
                        if bl.get("substeps") == 2 else
                        "k_collide_stream<double, false> (both wall bands, one launch per substep)"),
                algorithmic_bytes_per_launch=bl["bytes"], launch_ms=bl["ms"])
    cfg5 = None
    if not args.no_cfg5:
        print(f"bench: Re_tau=395 posts channel, fp32, {args.cfg5_steps} steps", file=sys.stderr, flush=True)
        cfg5 = run_cfg5(mod, args.cfg5_steps, 2)
        print(f"bench: Re_tau=395 {cfg5['glups']} GLUPS, band {cfg5['band_launch']['gbs']} GB/s, "
              f"spectra {cfg5['spectra_s_three_grids']} s", file=sys.stderr, flush=True)
    prec_report = None
    if not args.no_precision_report:
        print("bench: precision tier report (64x32x32 channel, 500 steps)", file=sys.stderr, flush=True)
        prec_report = precision_tier(mod)
    queue = None
    if not args.no_queue and args.queue_requests > 0:
        print(f"bench: slip queue ({args.queue_requests} requests at {args.queue_rate}/s)", file=sys.stderr, flush=True)
        queue = run_queue(mod, args.queue_requests, args.queue_rate, 2025, min(args.queue_check, args.queue_requests),
                          args.queue_envelope)
        print(f"bench: queue p50 {queue['p50_s']:.2f} s p99 {queue['p99_s']:.2f} s, {queue['glups']:.2f} GLUPS",
              file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        print(f"bench: cpu baseline ({args.cpu_steps} steps)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(mod, args.cpu_steps)
    line = {
        "metric": METRIC,
        "value": round(main_t["glups"], 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_t["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16-storage/f32"}[main_prec],
        "data": "synthetic (seeded laminar profile + 10% noise)",
        "config": {"workload": "Re_tau=180 superhydrophobic-stripe channel, coarse 512x256x128, "
                               "factor-2 bands 2x1024x512x17, 33.6M cells, constant body force",
                   "lu_per_step": main_t["lu_per_step"], "parallelism": "single GPU"},
        "hbm_roofline_frac_step": round(main_t["glups"] * BYTES_PER_LU[main_prec] / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64_flux_controller": flux,
        "slip_queue": queue,
        "precision_tier": prec_report,
        "retau395_posts_fp32": cfg5,
        "hbm_probe_gbs": probe,
        "tiers": {k: {"glups": round(v["glups"], 4), "ms_per_step": round(v["ms_per_step"], 4),
                      "hbm_frac": round(v["glups"] * BYTES_PER_LU[k] / HBM_PEAK_GBS, 4),
                      "kernels_ms": {a: round(b, 4) for a, b in v["kernels_ms"].items()}}
                  for k, v in res.items()},
    }
    if "fp64" in res and "fp32" in res:
        line["fp32_over_fp64"] = round(res["fp32"]["glups"] / res["fp64"]["glups"], 3)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
