"""ctypes binding of libdnsg.so (include/dnsg.h) for tests, bench.py and scripting.

The product is the C-ABI library built in-tree at build/libdnsg.so (HIP kernels
for gfx950) and the C driver build/dnsg-run.  This module only marshals
parameters and numpy arrays across that ABI; it computes nothing itself, and it
fails loudly when the library is missing instead of falling back to anything.

The package directory name contains hyphens, so load it by path:

    spec = importlib.util.spec_from_file_location("dnsg_amd", ".../__init__.py")
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libdnsg.so")
DRIVER_PATH = os.path.join(HERE, "build", "dnsg-run")

PATTERN = {"none": 0, "rslip": 1, "pslip": 2, "sslip": 3}
FORCE = {"flux": 0, "const": 1}
PREC = {"fp64": 0, "fp32": 1, "fp16": 2}
INIT = {"rest": 0, "laminar_noise": 1}
STATUS = {0: "ok", -1: "bad argument", -2: "HIP error", -3: "NaN in wall shear stress",
          -4: "out of memory", -5: "out of sequence", -6: "I/O error"}


class Params(ctypes.Structure):
    """dnsg_params (definitions.h names + runtime extensions)."""
    _fields_ = [
        ("Rebs", ctypes.c_double), ("ub", ctypes.c_double),
        ("zdv", ctypes.c_int), ("ydv", ctypes.c_int), ("xdv", ctypes.c_int),
        ("TSTR", ctypes.c_int), ("TEND", ctypes.c_int), ("CFL", ctypes.c_double),
        ("gratio", ctypes.c_int), ("zds", ctypes.c_int),
        ("gyplussize", ctypes.c_int), ("wyplussize", ctypes.c_int),
        ("gxplussize", ctypes.c_int), ("wxplussize", ctypes.c_int),
        ("pattern", ctypes.c_int), ("ITER", ctypes.c_long),
        ("force_mode", ctypes.c_int), ("g_const", ctypes.c_double), ("tau_c", ctypes.c_double),
        ("precision", ctypes.c_int), ("init_kind", ctypes.c_int), ("seed", ctypes.c_uint64),
        ("u_max", ctypes.c_double), ("noise", ctypes.c_double),
    ]


class Dims(ctypes.Structure):
    _fields_ = [
        ("Xc", ctypes.c_int), ("Yc", ctypes.c_int), ("zdc", ctypes.c_int),
        ("Xf", ctypes.c_int), ("Yf", ctypes.c_int), ("zdf", ctypes.c_int),
        ("GR", ctypes.c_int), ("NZs", ctypes.c_int), ("NZ", ctypes.c_int),
        ("utime", ctypes.c_long), ("up", ctypes.c_long), ("dt", ctypes.c_double),
        ("omega_c", ctypes.c_double), ("omega_f", ctypes.c_double),
        ("tau_c", ctypes.c_double), ("tau_f", ctypes.c_double),
        ("n_coarse", ctypes.c_size_t), ("n_band", ctypes.c_size_t), ("n_plane", ctypes.c_size_t),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


DIAG_FIELDS = ["time", "force", "momentum", "density", "flowrate", "str_up", "str_down", "stress", "pgrad"]


class Diag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in DIAG_FIELDS]


# exported entry points and their signatures (all return int status)
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "dnsg_params_default": [ctypes.POINTER(Params)],
    "dnsg_dims_from_params": [ctypes.POINTER(Params), ctypes.POINTER(Dims)],
    "dnsg_create": [ctypes.POINTER(Params), ctypes.POINTER(_P)],
    "dnsg_destroy": [_P],
    "dnsg_get_dims": [_P, ctypes.POINTER(Dims)],
    "dnsg_init_state": [_P],
    "dnsg_upload_state": [_P, _DP, _DP, _DP, _DP, _DP, ctypes.c_double],
    "dnsg_download_state": [_P, _DP, _DP, _DP, _DP, _DP, _DP],
    "dnsg_step": [_P, ctypes.c_int, ctypes.POINTER(Diag)],
    "dnsg_run_timed": [_P, ctypes.c_int, ctypes.POINTER(Diag), _DP],
    "dnsg_kernel_times": [_P, ctypes.c_int, _DP],
    "dnsg_hbm_probe": [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, _DP],
    "dnsg_encalc2": [_P, ctypes.c_long, ctypes.POINTER(Diag)],
    "dnsg_wshearstr": [_P, ctypes.c_long, ctypes.POINTER(Diag)],
    "dnsg_xcomputations": [_P],
    "dnsg_xfcomputations": [_P],
    "dnsg_ctfdtransfer": [_P, ctypes.c_int],
    "dnsg_ftcdtransfer": [_P],
    "dnsg_exchange": [_P],
    "dnsg_fexchange": [_P],
    "dnsg_mass": [_P, _DP],
    "dnsg_statistics": [_P, _DP],
    "dnsg_stats_accumulate": [_P],
    "dnsg_stats_average": [_P, _DP, ctypes.POINTER(ctypes.c_long)],
    "dnsg_stats_reset": [_P],
    "dnsg_velocity_snapshot": [_P, ctypes.c_int, _DP, _DP],
}

_lib = None


def load(path=LIB_PATH):
    """Load libdnsg.so; raise if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libdnsg.so not built at {path}: run __graft_entry__.build() "
                           "or `make -C dns-solvers-with-grid-embedding_amd`")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.dnsg_last_error.argtypes = [_P]
    lib.dnsg_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def make_params(**kw):
    """dnsg_params with the library defaults, overridden by keyword (names as in
    definitions.h; pattern/force_mode/precision/init_kind also accept strings)."""
    lib = load()
    p = Params()
    lib.dnsg_params_default(ctypes.byref(p))
    maps = {"pattern": PATTERN, "force_mode": FORCE, "precision": PREC, "init_kind": INIT}
    for k, v in kw.items():
        if k in maps and isinstance(v, str):
            v = maps[k][v]
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k}")
        setattr(p, k, type(getattr(p, k))(v))
    return p


def dims_of(params):
    d = Dims()
    rc = load().dnsg_dims_from_params(ctypes.byref(params), ctypes.byref(d))
    if rc:
        raise ValueError(f"invalid parameters ({STATUS.get(rc, rc)})")
    return d


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


class DnsgError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"{STATUS.get(rc, rc)}: {msg}")
        self.rc = rc


class Solver:
    """One device-resident solver (a dnsg_t handle)."""

    def __init__(self, params=None, **kw):
        self.lib = load()
        self.params = params if params is not None else make_params(**kw)
        self.h = ctypes.c_void_p()
        rc = self.lib.dnsg_create(ctypes.byref(self.params), ctypes.byref(self.h))
        if rc:
            msg = self.lib.dnsg_last_error(self.h).decode() if self.h else ""
            if self.h:
                self.lib.dnsg_destroy(self.h)
                self.h = ctypes.c_void_p()
            raise DnsgError(rc, msg)
        self.dims = Dims()
        self.lib.dnsg_get_dims(self.h, ctypes.byref(self.dims))

    def _check(self, rc):
        if rc:
            raise DnsgError(rc, self.lib.dnsg_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.dnsg_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # shapes of the reference file layouts
    def shapes(self):
        d = self.dims
        return dict(coarse=(d.Xc, d.Yc, d.zdc, 19), band=(d.Xf, d.Yf, d.zdf, 19), levels=(3, d.Xc, d.Yc, 23))

    def init_state(self):
        self._check(self.lib.dnsg_init_state(self.h))

    def upload(self, coarse=None, lower=None, upper=None, lin=None, uin=None, force=0.0):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in
                (coarse, lower, upper, lin, uin)]
        self._check(self.lib.dnsg_upload_state(self.h, *[_ptr(a) for a in arrs], float(force)))

    def download(self):
        s = self.shapes()
        c = np.empty(s["coarse"]); lo = np.empty(s["band"]); up = np.empty(s["band"])
        li = np.empty(s["levels"]); ui = np.empty(s["levels"]); f = np.zeros(1)
        self._check(self.lib.dnsg_download_state(self.h, _ptr(c), _ptr(lo), _ptr(up), _ptr(li), _ptr(ui), _ptr(f)))
        return dict(coarse=c, lower=lo, upper=up, lin=li, uin=ui, force=float(f[0]))

    def step(self, n=1):
        d = (Diag * max(n, 1))()
        self._check(self.lib.dnsg_step(self.h, n, d))
        return diag_array(d, n)

    def run_timed(self, n):
        d = (Diag * n)()
        ms = ctypes.c_double()
        self._check(self.lib.dnsg_run_timed(self.h, n, d, ctypes.byref(ms)))
        return ms.value, diag_array(d, n)

    def kernel_times(self, n):
        """average ms per launch: reductions, coarse CS, sampling, band CS, interface"""
        ms = np.zeros(5)
        self._check(self.lib.dnsg_kernel_times(self.h, n, _ptr(ms)))
        return dict(zip(("reduce", "coarse", "sample", "band", "interface"), ms))

    def mass(self):
        m = ctypes.c_double()
        self._check(self.lib.dnsg_mass(self.h, ctypes.byref(m)))
        return m.value

    def statistics(self):
        out = np.zeros((9, self.dims.NZ))
        self._check(self.lib.dnsg_statistics(self.h, _ptr(out)))
        return out

    def stats_accumulate(self):
        self._check(self.lib.dnsg_stats_accumulate(self.h))

    def stats_average(self):
        out = np.zeros((9, self.dims.NZ))
        cnt = ctypes.c_long()
        self._check(self.lib.dnsg_stats_average(self.h, _ptr(out), ctypes.byref(cnt)))
        return out, cnt.value

    def stats_reset(self):
        self._check(self.lib.dnsg_stats_reset(self.h))

    def velocity(self, grid):
        d = self.dims
        X, Y, Z = (d.Xc, d.Yc, d.zdc) if grid == 0 else (d.Xf, d.Yf, d.zdf)
        vel = np.empty((X, Y, Z, 3)); den = np.empty((X, Y, Z))
        self._check(self.lib.dnsg_velocity_snapshot(self.h, grid, _ptr(vel), _ptr(den)))
        return vel, den

    # reference-granular entry points
    def encalc2(self, ss):
        d = Diag()
        self._check(self.lib.dnsg_encalc2(self.h, ss, ctypes.byref(d)))
        return d

    def wshearstr(self, ss):
        d = Diag()
        self._check(self.lib.dnsg_wshearstr(self.h, ss, ctypes.byref(d)))
        return d

    def xcomputations(self):
        self._check(self.lib.dnsg_xcomputations(self.h))

    def xfcomputations(self):
        self._check(self.lib.dnsg_xfcomputations(self.h))

    def ctfdtransfer(self, it):
        self._check(self.lib.dnsg_ctfdtransfer(self.h, it))

    def ftcdtransfer(self):
        self._check(self.lib.dnsg_ftcdtransfer(self.h))

    def exchange(self):
        self._check(self.lib.dnsg_exchange(self.h))

    def fexchange(self):
        self._check(self.lib.dnsg_fexchange(self.h))


def hbm_probe(nbytes=2 << 30, kind=0, reps=5):
    """achievable HBM bandwidth (GB/s): kind 0 copy, 1 read, 2 write"""
    g = ctypes.c_double()
    rc = load().dnsg_hbm_probe(nbytes, kind, reps, ctypes.byref(g))
    if rc:
        raise DnsgError(rc, "hbm probe")
    return g.value


def diag_array(d, n):
    return {f: np.array([getattr(d[i], f) for i in range(n)]) for f in DIAG_FIELDS}
