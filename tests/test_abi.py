"""CPU checks of the C-ABI library: it loads without a GPU, exports every entry
point include/dnsg.h declares, and derives sizes and relaxation rates exactly as
the oracle (and hence the reference, main:139-155, 432-469) does."""
import ctypes
import re

import pytest

from conftest import REPO
from helpers import Oracle, dnsg

HEADER = f"{REPO}/include/dnsg.h"


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(dnsg_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = dnsg.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), f"{n} declared in dnsg.h but not exported"
    special = {"dnsg_last_error", "dnsg_queue_last_error", "dnsg_queue_clock", "dnsg_solve_bytes"}  # non-int returns
    assert set(names) - special == set(dnsg.SIGNATURES)


def test_params_struct_matches_header_layout():
    # field order/size of the ctypes mirror must match the C struct
    assert ctypes.sizeof(dnsg.Params) == 152
    assert ctypes.sizeof(dnsg.Diag) == 9 * 8
    assert ctypes.sizeof(dnsg.QueueOpts) == 48
    assert ctypes.sizeof(dnsg.SlipResult) == 72


CASES = [
    dict(xdv=64, ydv=32, zdv=32, gratio=2, zds=8, Rebs=182.5, ub=0.0333333333333, CFL=4.0),
    dict(xdv=16, ydv=8, zdv=12, gratio=3, zds=5, Rebs=20.0, ub=0.0333333333333, CFL=5.0),
    dict(xdv=512, ydv=256, zdv=128, gratio=2, zds=16, tau_c=0.5058541666666667, CFL=0.1),
]


@pytest.mark.parametrize("kw", CASES)
def test_dims_match_oracle(kw, oracle_built):
    p = dnsg.make_params(**kw)
    d = dnsg.dims_of(p)
    o = Oracle(p)
    for k, _ in dnsg.Dims._fields_:
        assert getattr(d, k) == getattr(o.dims, k), k


def test_invalid_parameters_are_rejected():
    for kw in [dict(gratio=0), dict(zds=2), dict(pattern=1, gyplussize=5, wyplussize=4),
               dict(CFL=1e9), dict(precision=7), dict(tau_c=0.4)]:
        with pytest.raises(ValueError):
            dnsg.dims_of(dnsg.make_params(**kw))
