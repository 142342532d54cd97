"""Pin the oracle to the reference solver.

The fixtures in tests/golden/*.json were written by tests/golden/make_golden.py,
which runs the reference (compiled from its own sources) on one MPI rank.  Each
case feeds both programs the same '-p' restart files (or the reference's '-s'
rest start) and compares, byte for byte, every checkpoint the reference wrote
(both f/ftemp halves of lbf-c/slf/suf, the three interface time levels,
force.NNNN), the last velocity/density snapshots, the statistics files, the
output_twoD2 profile and the six per-step text logs.  The '-i' cases start both
from the same synthetic uniform-grid file and compare the initializer's output:
the step-0 checkpoint, snapshots and statistics, the step-0 log lines and the
start mass.
"""
import glob
import hashlib
import json
import os
import re
import subprocess

import pytest

from conftest import GOLDEN, ORC_CLI

CASES = sorted(c for c in glob.glob(os.path.join(GOLDEN, "*.json"))
               if json.load(open(c)).get("kind") not in ("budget", "masks"))
BUDGET_CASES = sorted(c for c in glob.glob(os.path.join(GOLDEN, "*.json")) if json.load(open(c)).get("kind") == "budget")


def _sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@pytest.mark.parametrize("case", CASES, ids=[os.path.basename(c)[:-5] for c in CASES])
def test_oracle_matches_reference(case, tmp_path, oracle_built):
    fx = json.load(open(case))
    args = [f"{k}={v}" for k, v in fx["params"].items()]
    if fx["mode"] == "p":
        args += [f"{k}={v}" for k, v in fx["init"].items()]
    elif fx["mode"] == "i":
        uni = args + [f"{k}={v}" for k, v in fx["init"].items()]
        subprocess.run([ORC_CLI, "uniform", str(tmp_path / "lbf.0000")] + uni, check=True)
        args += ["init_kind=2"]
    else:
        args += ["init_kind=0"]
    subprocess.run([ORC_CLI, "init", str(tmp_path)] + args, check=True)
    out = subprocess.run([ORC_CLI, "run", str(tmp_path)] + args + ["threads=4"], check=True,
                         capture_output=True, text=True).stdout
    for name, digest in fx["sha256"].items():
        assert _sha(tmp_path / name) == digest, f"{name} differs from the reference"
    for name, text in fx["logs"].items():
        got = (tmp_path / name).read_text()
        if fx["mode"] == "i":  # the step-0 line only
            got = got.splitlines(keepends=True)[0]
        assert got == text, f"{name} differs from the reference"
    ref_mass = [re.search(r":\s*([-0-9.eE+]+)", l).group(1) for l in fx["mass"]]
    got_mass = [re.search(r":\s*([-0-9.eE+]+)", l).group(1) for l in out.splitlines() if "Average" in l]
    got_mass = got_mass[: len(ref_mass)]
    assert got_mass == ref_mass


@pytest.mark.parametrize("case", BUDGET_CASES, ids=[os.path.basename(c)[:-5] for c in BUDGET_CASES])
def test_budget_oracle_matches_reference_tool(case, tmp_path, oracle_built):
    """the TKE-budget restatement against the reference's budget tool run on the same
    snapshots (TKE-budget/computations): every file it writes, byte for byte"""
    fx = json.load(open(case))
    args = [f"{k}={v}" for k, v in fx["params"].items()] + [f"{k}={v}" for k, v in fx["init"].items()]
    subprocess.run([ORC_CLI, "init", str(tmp_path)] + args, check=True)
    subprocess.run([ORC_CLI, "run", str(tmp_path)] + args + ["threads=4"], check=True, capture_output=True)
    post = tmp_path / "post"
    post.mkdir()
    subprocess.run([ORC_CLI, "budget", str(post)] + args + [f"TSTR={fx['TSTR']}", f"TEND={fx['TEND']}"],
                   check=True)
    assert sorted(os.listdir(post)) == sorted(fx["sha256"])
    for name, digest in fx["sha256"].items():
        assert _sha(post / name) == digest, f"{name} differs from the reference tool's"
