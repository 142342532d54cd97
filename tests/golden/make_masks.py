#!/usr/bin/env python3
"""Write tests/golden/masks.json: the reference's own node-code arrays V.s, V.sls,
V.sus (solid_init.c + slip_*_init.c, one rank) for a set of geometries, as SHA-256
digests of their int32 bytes and the count of every code.

Each case builds oracle/mask_ref_main.c against the reference's sources with
oracle/build_mask_ref.sh (in this container only; /root/reference is not on the GPU
side) and runs it under MPICH's mpiexec.  The fixture is data: digests and counts.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BUILD = os.path.join(REPO, "oracle", "build_mask_ref.sh")
OUT = os.path.join(REPO, "tests", "golden", "masks.json")
PAT = {"none": 0, "rslip": 1, "pslip": 2, "sslip": 3}

MICRO = dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4)
CASES = {
    "micro_none": dict(MICRO, pattern="none"),
    "micro_rslip": dict(MICRO, pattern="rslip", gyplussize=4, wyplussize=4),
    "micro_pslip": dict(MICRO, pattern="pslip", gxplussize=4, wxplussize=4, gyplussize=4, wyplussize=4),
    "micro_sslip": dict(MICRO, pattern="sslip", gxplussize=6, wxplussize=2),
    "gr3_rslip": dict(MICRO, gratio=3, zds=5, pattern="rslip", gyplussize=6, wyplussize=6),
    "gr1_rslip": dict(MICRO, gratio=1, pattern="rslip", gyplussize=4, wyplussize=4),
    "ragged_pslip": dict(xdv=13, ydv=7, zdv=9, gratio=2, zds=5, pattern="pslip", gxplussize=2, wxplussize=24,
                         gyplussize=5, wyplussize=2),
    "wide_sslip": dict(xdv=256, ydv=6, zdv=9, gratio=2, zds=4, pattern="sslip", gxplussize=3, wxplussize=5),
    "cfg1_none": dict(xdv=64, ydv=32, zdv=32, gratio=2, zds=8, pattern="none"),
    # the Re_tau=180 stripe channel and the Re_tau=395 square-post channel at full size
    "retau180_rslip": dict(xdv=512, ydv=256, zdv=128, gratio=2, zds=16, pattern="rslip", gyplussize=32,
                           wyplussize=32),
    "retau395_pslip": dict(xdv=1024, ydv=512, zdv=256, gratio=2, zds=32, pattern="pslip", gxplussize=32,
                           wxplussize=32, gyplussize=32, wyplussize=32),
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def counts(a):
    v, c = np.unique(a, return_counts=True)
    return {str(int(k)): int(n) for k, n in zip(v, c)}


def run(name, cfg):
    args = [f"{k}={PAT[v] if k == 'pattern' else v}" for k, v in cfg.items()]
    subprocess.run([BUILD, "mask_" + name] + args, check=True, stdout=subprocess.DEVNULL)
    exe = os.path.join(REPO, "oracle", "_ref", "mask_" + name, "mask-dump")
    with tempfile.TemporaryDirectory(prefix="dnsg_masks_") as work:
        out = os.path.join(work, "masks.bin")
        env = dict(os.environ, LD_LIBRARY_PATH="/opt/conda/lib")
        subprocess.run(["/opt/conda/bin/mpiexec", "-n", "1", exe, out], cwd=work, env=env, check=True)
        a = np.fromfile(out, dtype="<i4")
    xdc, ydc, zdc, xdf, ydf, zdf = (int(v) for v in a[:6])
    nc, nf = xdc * ydc * zdc, xdf * ydf * zdf
    s = a[6:6 + nc].reshape(xdc, ydc, zdc)
    sl = a[6 + nc:6 + nc + nf].reshape(xdf, ydf, zdf)
    su = a[6 + nc + nf:6 + nc + 2 * nf].reshape(xdf, ydf, zdf)
    assert a.size == 6 + nc + 2 * nf
    return dict(params=cfg, dims=[xdc, ydc, zdc, xdf, ydf, zdf],
                sha256={"s": digest(s), "sls": digest(sl), "sus": digest(su)},
                counts={"s": counts(s), "sls": counts(sl), "sus": counts(su)},
                wall_lower_gas_fraction=float(np.mean(sl[:-1, :-1, 0] == -2)),
                wall_upper_gas_fraction=float(np.mean(su[:-1, :-1, -1] == -2)))


def main():
    names = sys.argv[1:] or list(CASES)
    fx = json.load(open(OUT)) if os.path.exists(OUT) else {"kind": "masks", "cases": {}}
    for n in names:
        print(f"masks: {n}", flush=True)
        fx["cases"][n] = run(n, CASES[n])
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
