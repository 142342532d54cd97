"""Node codes (solid/gas masks and stripe/post maps) bit for bit against the
reference's own arrays.

tests/golden/masks.json holds, per geometry, the SHA-256 digests of the int32 arrays
V.s, V.sls and V.sus that the reference's solid_init.c and slip_*_init.c build on one
rank (dumped by oracle/mask_ref_main.c linked with those sources; written by
tests/golden/make_masks.py).  The oracle's restatement (orc_ref_masks) is checked
against them here on the CPU; the GPU handle's export (dnsg_node_codes, read back
from the device wall maps the kernels use) is checked against them in the gpu tests,
at the full Re_tau=180 and Re_tau=395 sizes too.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import Oracle, dnsg, oracle_lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "masks.json")
FIX = json.load(open(GOLDEN))["cases"]
SMALL = [k for k in FIX if not k.startswith("retau")]
REF = dict(Rebs=20.0, ub=0.0333333333333, CFL=1.0)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def params(name, **extra):
    return dnsg.make_params(**FIX[name]["params"], **REF, **extra)


def oracle_masks(p):
    o = Oracle(p)
    d = o.dims
    sc = np.empty((d.Xc + 1, d.Yc + 1, d.zdc), dtype=np.int32)
    sl = np.empty((d.Xf + 1, d.Yf + 1, d.zdf), dtype=np.int32)
    su = np.empty_like(sl)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    oracle_lib().orc_ref_masks(o.o, P(sc), P(sl), P(su))
    o.close()
    return sc, sl, su


@pytest.mark.parametrize("name", SMALL)
def test_oracle_masks_match_reference(name):
    sc, sl, su = oracle_masks(params(name))
    fx = FIX[name]
    assert list(sc.shape) == fx["dims"][:3] and list(sl.shape) == fx["dims"][3:]
    assert digest(sc) == fx["sha256"]["s"]
    assert digest(sl) == fx["sha256"]["sls"]
    assert digest(su) == fx["sha256"]["sus"]


def test_fixture_closed_forms():
    """what the digests encode, restated from the definitions: the Re_tau=395 posts
    (32 x 32 fine cells in a 64 x 64 period) leave a solid fraction of exactly 1/4 on
    each wall, the Re_tau=180 ridges (32 gas in a period of 64) a gas fraction of 1/2"""
    assert FIX["retau395_pslip"]["wall_lower_gas_fraction"] == 0.75
    assert FIX["retau395_pslip"]["wall_upper_gas_fraction"] == 0.75
    assert FIX["retau180_rslip"]["wall_lower_gas_fraction"] == 0.5
    for name, fx in FIX.items():
        xdc, ydc, zdc, xdf, ydf, zdf = fx["dims"]
        assert sum(fx["counts"]["s"].values()) == xdc * ydc * zdc
        # the coarse grid holds only fluid (0) and periodic/bounded (1) nodes
        assert set(fx["counts"]["s"]) <= {"0", "1"}
        assert fx["counts"]["s"]["0"] == (xdc - 2) * (ydc - 2) * (zdc - 2)
