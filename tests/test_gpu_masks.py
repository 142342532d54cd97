"""The GPU handle's node codes (dnsg_node_codes: the device wall maps the
collide-stream reads, expanded with the boundary set-up of its launches) bit for
bit against the reference's own V.s, V.sls, V.sus (tests/golden/masks.json, dumped
from the reference's solid_init.c and slip_*_init.c), for every geometry of the
fixture: the parity geometries, the 64x32x32 channel and the full-size Re_tau=180
stripe and Re_tau=395 square-post channels."""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import dnsg

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "masks.json")
FIX = json.load(open(GOLDEN))["cases"]
REF = dict(Rebs=20.0, ub=0.0333333333333, CFL=1.0)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("name", list(FIX))
def test_node_codes_match_reference(name):
    fx = FIX[name]
    prec = "fp32" if name.startswith("retau") else "fp64"
    print(f"node codes: {name} {fx['params']} ({prec})", flush=True)
    s = dnsg.Solver(dnsg.make_params(**fx["params"], **REF, precision=prec, force_mode="const", g_const=1e-6))
    sc, sl, su = s.node_codes()
    s.close()
    assert list(sc.shape) == fx["dims"][:3] and list(sl.shape) == fx["dims"][3:]
    assert digest(sc) == fx["sha256"]["s"]
    assert digest(sl) == fx["sha256"]["sls"]
    assert digest(su) == fx["sha256"]["sus"]
    if name == "retau395_pslip":  # solid posts cover exactly a quarter of each wall
        assert np.mean(sl[:-1, :-1, 0] == -1) == 0.25
        assert np.mean(su[:-1, :-1, -1] == -1) == 0.25
