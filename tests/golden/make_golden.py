#!/usr/bin/env python3
"""Write the golden fixtures that pin the oracle to the reference solver.

For every case below this script
  1. builds the reference from its own sources for that parameter set
     (oracle/build_ref.sh -> oracle/_ref/<case>/run-dns-mgrd),
  2. writes the seeded initial state as '-p' restart files with the oracle CLI
     (or uses the reference's own '-s' rest start, or, for its '-i' start, writes
     a synthetic uniform-grid file lbf.0000 with the oracle CLI),
  3. runs the reference on one MPI rank,
  4. stores SHA-256 digests of every checkpoint file it wrote, the six per-step
     text logs, the final turb-fld statistics and 2dresults-lb.bdat in
     tests/golden/<case>.json.

The BUDGET cases pin the TKE-budget oracle (oracle/budget_oracle.c) the same way:
the oracle's solver writes a few velocity/density snapshots, the reference's
budget tool (TKE-budget/computations, built by oracle/build_budget_ref.sh) runs on
them from a subdirectory, as its README prescribes, and the digests of every file
it writes are stored.

It runs only where the reference sources and MPICH are present; the fixtures
it writes are plain data (digests, text logs, numbers) and are what the CPU
tests compare the oracle against.
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
ORC_CLI = os.path.join(REPO, "oracle", "build", "orc_cli")
BUILD_REF = os.path.join(REPO, "oracle", "build_ref.sh")
BUILD_BUDGET = os.path.join(REPO, "oracle", "build_budget_ref.sh")

MICRO = dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4, Rebs="20.0", ub="0.0333333333333",
             CFL="5.0", ITER=50)
CASES = {
    # seeded laminar + noise through '-p' restart files, no surface pattern
    "micro_none": dict(MICRO),
    # streamwise ridges, posts and transverse ridges on both walls
    "micro_rslip": dict(MICRO, pattern=1, gyplussize=4, wyplussize=4),
    "micro_pslip": dict(MICRO, pattern=2, gxplussize=4, wxplussize=4, gyplussize=4, wyplussize=4),
    "micro_sslip": dict(MICRO, pattern=3, gxplussize=6, wxplussize=2),
    # grid ratio 3 (two intermediate substeps, t = 1/3, 2/3)
    "micro_gr3": dict(MICRO, gratio=3, zds=5, CFL="5.0", pattern=1, gyplussize=6, wyplussize=6),
    # the reference's own '-s' start: rest state, impulsive flux-controlled start
    "micro_rest": dict(MICRO, pattern=1, gyplussize=4, wyplussize=4, mode="s"),
    # FLTR: box-filtered fine->coarse injection (ftcdtransfer_lesfltr.c), grid ratio 4;
    # the 5-layer filter slab reaches both wall layers
    "micro_fltr": dict(MICRO, gratio=4, zds=6, pattern=1, gyplussize=4, wyplussize=4, fltr=1),
    # the '-i' start from a uniform-grid field (initializer.c, inter_init.c): one
    # uniform grid finer than the fine grid (physically consistent extents) and one
    # coarser than it, whose first lower-band layer takes the one-sided z stencil.
    # Only the initializer's output is recorded (the .0000 checkpoint, the step-0 log
    # lines and the start mass): the step after it reads the older interface levels'
    # ghost cells, which the reference's level copy leaves unlike their periodic
    # images (inter_init.c:245-257)
    "micro_uniform_fine": dict(MICRO, pattern=1, gyplussize=4, wyplussize=4, mode="i", ITER=1,
                               xdu=64, ydu=32, zdu=50),
    "micro_uniform_coarse": dict(MICRO, mode="i", ITER=1, xdu=24, ydu=12, zdu=13),
    # the 64x32x32 channel with factor-2 bands of 8 layers, 200 steps
    "cfg1_200": dict(xdv=64, ydv=32, zdv=32, gratio=2, zds=8, Rebs="182.5",
                     ub="0.0333333333333", CFL="4.0", ITER=200),
}
INIT = dict(seed=42, u_max="0.05", noise="0.05")
# snapshots TSTR..TEND of a short solver run, then the budget tool over them: the
# thinnest bands (the one-sided z stencils reach into the neighbouring rows there)
# with the fluctuation dump of snapshot 0, and deeper bands at grid ratio 3
BUDGET = {
    "budget_micro": dict(solver=dict(MICRO, pattern=1, gyplussize=4, wyplussize=4, ITER=3), TSTR=0, TEND=2),
    "budget_gr3": dict(solver=dict(xdv=12, ydv=8, zdv=14, gratio=3, zds=8, Rebs="20.0", ub="0.0333333333333",
                                   CFL="5.0", ITER=3, pattern=2, gxplussize=6, wxplussize=6, gyplussize=6,
                                   wyplussize=6), TSTR=1, TEND=3),
}
BUDGET_KEYS = ["Rebs", "ub", "xdv", "ydv", "zdv", "CFL", "gratio", "zds"]
LOGS = ["flowrate.txt", "density.txt", "str-up.txt", "str-do.txt", "w-str.txt", "p-grad.txt"]
REF_KEYS = ["Rebs", "ub", "xdv", "ydv", "zdv", "TSTR", "TEND", "CFL", "gratio", "zds",
            "gyplussize", "wyplussize", "gxplussize", "wxplussize", "pattern", "ITER", "fltr",
            "xdu", "ydu", "zdu"]
UNIFORM = dict(seed=77, u_max="0.05", noise="0.1")


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def run_case(name, cfg):
    cfg = dict(cfg)
    mode = cfg.pop("mode", "p")
    ref_args = [f"{k}={v}" for k, v in cfg.items() if k in REF_KEYS]
    subprocess.run([BUILD_REF, name] + ref_args, check=True, stdout=subprocess.DEVNULL)
    exe = os.path.join(REPO, "oracle", "_ref", name, "run-dns-mgrd")
    work = tempfile.mkdtemp(prefix=f"dnsg_golden_{name}_")
    try:
        if mode == "p":
            init_args = ref_args + [f"{k}={v}" for k, v in INIT.items()]
            subprocess.run([ORC_CLI, "init", work] + init_args, check=True)
        if mode == "i":
            uni_args = ref_args + [f"{k}={v}" for k, v in UNIFORM.items()]
            subprocess.run([ORC_CLI, "uniform", os.path.join(work, "lbf.0000")] + uni_args, check=True)
        env = dict(os.environ, LD_LIBRARY_PATH="/opt/conda/lib")
        out = subprocess.run(["/opt/conda/bin/mpiexec", "-n", "1", exe, "-" + mode], cwd=work,
                             env=env, check=True, capture_output=True, text=True).stdout
        fixture = dict(params=cfg, mode=mode, init=INIT if mode == "p" else None, sha256={}, logs={},
                       mass=[l.strip() for l in out.splitlines() if "Average Density" in l])
        files = sorted(os.listdir(work))
        if mode == "i":
            fixture["init"] = UNIFORM
            fixture["mass"] = fixture["mass"][:1]
            for f in files:
                if f.endswith(".0000") and f != "lbf.0000" or f.startswith("turb-fld.") and f.endswith(".000"):
                    fixture["sha256"][f] = sha(os.path.join(work, f))
            for f in LOGS:
                with open(os.path.join(work, f)) as fh:
                    fixture["logs"][f] = fh.readline()
            with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
                json.dump(fixture, fh, indent=1, sort_keys=True)
            print(f"{name}: {len(fixture['sha256'])} digests of the initializer's output")
            return
        for f in files:
            if f.startswith(("lbf-", "force.")) and not f.endswith(".0000"):
                fixture["sha256"][f] = sha(os.path.join(work, f))
        turb = [f for f in files if f.startswith("turb-fld.")]
        last = max(f.split(".")[-1] for f in files if f.startswith("vel-c."))
        snaps = [f"{v}-{g}.{last}" for v in ("vel", "den") for g in ("c", "slf", "suf")]
        for f in turb[-2:] + ["2dresults-lb.bdat"] + snaps:
            fixture["sha256"][f] = sha(os.path.join(work, f))
        for f in LOGS:
            with open(os.path.join(work, f)) as fh:
                fixture["logs"][f] = fh.read()
        with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
            json.dump(fixture, fh, indent=1, sort_keys=True)
        print(f"{name}: {len(fixture['sha256'])} digests, {len(LOGS)} logs")
    finally:
        shutil.rmtree(work)


def run_budget_case(name, cfg):
    solver = cfg["solver"]
    args = [f"{k}={v}" for k, v in solver.items()] + [f"{k}={v}" for k, v in INIT.items()]
    tool = [f"{k}={solver[k]}" for k in BUDGET_KEYS] + [f"TSTR={cfg['TSTR']}", f"TEND={cfg['TEND']}"]
    subprocess.run([BUILD_BUDGET, name] + tool, check=True, stdout=subprocess.DEVNULL)
    exe = os.path.join(REPO, "oracle", "_ref", name, "compute-budget")
    work = tempfile.mkdtemp(prefix=f"dnsg_golden_{name}_")
    try:
        subprocess.run([ORC_CLI, "init", work] + args, check=True)
        subprocess.run([ORC_CLI, "run", work] + args + ["threads=4"], check=True, capture_output=True)
        post = os.path.join(work, "post")
        os.mkdir(post)
        env = dict(os.environ, LD_LIBRARY_PATH="/opt/conda/lib")
        subprocess.run(["/opt/conda/bin/mpiexec", "-n", "1", exe], cwd=post, env=env, check=True,
                       capture_output=True)
        fixture = dict(kind="budget", params=solver, init=INIT, TSTR=cfg["TSTR"], TEND=cfg["TEND"],
                       sha256={f: sha(os.path.join(post, f)) for f in sorted(os.listdir(post))})
        with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
            json.dump(fixture, fh, indent=1, sort_keys=True)
        print(f"{name}: {len(fixture['sha256'])} budget-tool files")
    finally:
        shutil.rmtree(work)


if __name__ == "__main__":
    if not os.path.exists(ORC_CLI):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True)
    wanted = sys.argv[1:] or list(CASES) + list(BUDGET)
    for n in wanted:
        if n in BUDGET:
            run_budget_case(n, BUDGET[n])
        else:
            run_case(n, CASES[n])
