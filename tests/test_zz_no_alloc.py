"""dnsg_step_noalloc (no counterpart in the reference; for callers that cannot
allocate) against dnsg_step: two calls in a row through the C-ABI, the second on a
smaller channel with fewer steps, each on a twin handle advanced by dnsg_step from
the same state.  Bar: the states (every lattice and interface level) and every
record bitwise equal, as both run the same kernels; n beyond the fixed array's
capacity is refused."""
import ctypes

import numpy as np
import pytest

from helpers import dnsg

pytestmark = pytest.mark.gpu

BIG = dict(xdv=64, ydv=32, zdv=16, gratio=2, zds=6, Rebs=40.0, ub=0.0333333333333, CFL=5.0,
           pattern="pslip", gxplussize=8, wxplussize=8, gyplussize=8, wyplussize=8)
SMALL = dict(xdv=16, ydv=8, zdv=12, gratio=3, zds=5, Rebs=20.0, ub=0.0333333333333, CFL=5.0,
             pattern="rslip", gyplussize=6, wyplussize=6)
INIT = dict(init_kind="laminar_noise", seed=5, u_max=0.05, noise=0.05)


def twins(kw):
    p = dnsg.make_params(**kw, **INIT)
    a, b = dnsg.Solver(p), dnsg.Solver(p)
    a.init_state()
    b.init_state()
    return a, b


@pytest.mark.parametrize("force_mode", ["flux", "const"])
def test_noalloc_step_equals_step(force_mode):
    extra = {} if force_mode == "flux" else dict(force_mode="const", g_const=2e-5)
    for kw, n in ((dict(BIG, **extra), 40), (dict(SMALL, **extra), 9)):
        print(f"noalloc vs step: {kw['xdv']}x{kw['ydv']}x{kw['zdv']} gratio {kw['gratio']}, "
              f"{force_mode}, {n} steps", flush=True)
        a, b = twins(kw)
        da = a.step(n)
        db = b.step_noalloc(n)
        for f in dnsg.DIAG_FIELDS:
            assert np.array_equal(da[f], db[f]), f
        ga, gb = a.download(), b.download()
        for k in ("coarse", "lower", "upper", "lin", "uin"):
            assert np.array_equal(ga[k], gb[k]), k
        assert ga["force"] == gb["force"]
        a.close()
        b.close()


def test_noalloc_rejects_too_many_steps():
    a, _ = twins(SMALL)
    n = dnsg.NOALLOC_MAX_STEPS + 1
    d = (dnsg.Diag * n)()
    rc = a.lib.dnsg_step_noalloc(a.h, n, d)
    assert rc == -1
    assert "DNSG_NOALLOC_MAX_STEPS" in a.lib.dnsg_last_error(a.h).decode()
    assert a.lib.dnsg_step_noalloc(a.h, 0, d) == 0
