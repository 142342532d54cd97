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
        ("u_max", ctypes.c_double), ("noise", ctypes.c_double), ("fltr", ctypes.c_int),
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


NOALLOC_MAX_STEPS = 4096  # DNSG_NOALLOC_MAX_STEPS of include/dnsg.h
DIAG_FIELDS = ["time", "force", "momentum", "density", "flowrate", "str_up", "str_down", "stress", "pgrad"]


class Diag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in DIAG_FIELDS]


# exported entry points and their signatures (all return int status)
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
class QueueOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_active", ctypes.c_int), ("check_every", ctypes.c_int),
                ("window", ctypes.c_int), ("tol", ctypes.c_double), ("max_steps", ctypes.c_long),
                ("floor_tol", ctypes.c_double)]


class SlipResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("precision", ctypes.c_int),
                ("steps", ctypes.c_long), ("ubulk", ctypes.c_double), ("u_noslip", ctypes.c_double),
                ("slip_length", ctypes.c_double), ("t_arrival", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES = {
    "dnsg_params_default": [ctypes.POINTER(Params)],
    "dnsg_dims_from_params": [ctypes.POINTER(Params), ctypes.POINTER(Dims)],
    "dnsg_create": [ctypes.POINTER(Params), ctypes.POINTER(_P)],
    "dnsg_destroy": [_P],
    "dnsg_get_dims": [_P, ctypes.POINTER(Dims)],
    "dnsg_node_codes": [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "dnsg_init_state": [_P],
    "dnsg_upload_state": [_P, _DP, _DP, _DP, _DP, _DP, ctypes.c_double],
    "dnsg_init_uniform": [_P, _DP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "dnsg_budget2d": [_P, _DP, _DP, _DP, _DP, _DP],
    "dnsg_budget2d_accumulate": [_P, _DP, _DP, _DP],
    "dnsg_budget2d_average": [_P, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_long)],
    "dnsg_budget2d_reset": [_P],
    "dnsg_download_state": [_P, _DP, _DP, _DP, _DP, _DP, _DP],
    "dnsg_upload_state_ghosted": [_P, _DP, _DP, _DP, ctypes.POINTER(_DP), ctypes.POINTER(_DP), ctypes.c_double],
    "dnsg_download_state_ghosted": [_P, _DP, _DP, _DP, ctypes.POINTER(_DP), ctypes.POINTER(_DP), _DP],
    "dnsg_step": [_P, ctypes.c_int, ctypes.POINTER(Diag)],
    "dnsg_step_noalloc": [_P, ctypes.c_int, ctypes.POINTER(Diag)],
    "dnsg_run_timed": [_P, ctypes.c_int, ctypes.POINTER(Diag), _DP],
    "dnsg_run_timed_sections": [_P, ctypes.c_int, ctypes.POINTER(Diag), _DP, _DP, ctypes.c_int],
    "dnsg_kernel_times": [_P, ctypes.c_int, _DP],
    "dnsg_band_launch_substeps": [_P, ctypes.POINTER(ctypes.c_int)],
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
    "dnsg_spectra": [_P, ctypes.c_int, _DP, _DP],
    "dnsg_queue_opts_default": [ctypes.POINTER(QueueOpts)],
    "dnsg_queue_create": [ctypes.POINTER(QueueOpts), ctypes.POINTER(_P)],
    "dnsg_queue_destroy": [_P],
    "dnsg_queue_submit": [_P, ctypes.POINTER(Params), ctypes.c_uint64, ctypes.c_double],
    "dnsg_queue_advance": [_P, ctypes.c_int, ctypes.POINTER(SlipResult), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "dnsg_queue_graph_launches": [_P, ctypes.POINTER(ctypes.c_long)],
    "dnsg_queue_info": [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(ctypes.c_size_t)],
}

_lib = None

# kernel groups of dnsg_kernel_times / dnsg_run_timed_sections, in their order
SECTIONS = ("reduce", "coarse", "sample", "band", "interface")


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
    lib.dnsg_queue_last_error.argtypes = [_P]
    lib.dnsg_queue_last_error.restype = ctypes.c_char_p
    lib.dnsg_queue_clock.argtypes = [_P]
    lib.dnsg_queue_clock.restype = ctypes.c_double
    lib.dnsg_solve_bytes.argtypes = [ctypes.POINTER(Params)]
    lib.dnsg_solve_bytes.restype = ctypes.c_size_t
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

    def node_codes(self):
        """V.s [Xc+1][Yc+1][zdc], V.sls and V.sus [Xf+1][Yf+1][zdf] (int32 node codes)"""
        d = self.dims
        sc = np.empty((d.Xc + 1, d.Yc + 1, d.zdc), dtype=np.int32)
        sl = np.empty((d.Xf + 1, d.Yf + 1, d.zdf), dtype=np.int32)
        su = np.empty_like(sl)
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        self._check(self.lib.dnsg_node_codes(self.h, P(sc), P(sl), P(su)))
        return sc, sl, su

    def init_state(self):
        self._check(self.lib.dnsg_init_state(self.h))

    def upload(self, coarse=None, lower=None, upper=None, lin=None, uin=None, force=0.0):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in
                (coarse, lower, upper, lin, uin)]
        self._check(self.lib.dnsg_upload_state(self.h, *[_ptr(a) for a in arrs], float(force)))

    def init_uniform(self, uniform, xdu, ydu, zdu):
        """'-i' start from the ftemp half of a uniform-grid lbf file, [xdu][ydu][zdu+2][19]"""
        u = np.ascontiguousarray(uniform, dtype=np.float64)
        if u.size != xdu * ydu * (zdu + 2) * 19:
            raise ValueError("uniform field size does not match xdu, ydu, zdu")
        self._check(self.lib.dnsg_init_uniform(self.h, _ptr(u), xdu, ydu, zdu))

    def budget2d(self, planes=False):
        """the TKE-budget tool's per-snapshot profiles of the current state:
        (turb [19][NZ], vort [3][NZ], ke [30][NZ], (ufl, vfl) [Xf][Yf] or None)"""
        NZ = self.dims.NZ
        turb, vort, ke = np.empty((19, NZ)), np.empty((3, NZ)), np.empty((30, NZ))
        uv = (np.empty((self.dims.Xf, self.dims.Yf)), np.empty((self.dims.Xf, self.dims.Yf))) if planes else None
        self._check(self.lib.dnsg_budget2d(self.h, _ptr(turb), _ptr(vort), _ptr(ke), _ptr(uv[0]) if uv else None,
                                           _ptr(uv[1]) if uv else None))
        return turb, vort, ke, uv

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

    def step_noalloc(self, n=1):
        """dnsg_step_noalloc: the same steps with no device or host allocation in the
        library (records through its fixed device array; n <= NOALLOC_MAX_STEPS)"""
        d = (Diag * max(n, 1))()
        self._check(self.lib.dnsg_step_noalloc(self.h, n, d))
        return diag_array(d, n)

    def budget2d_accumulate(self):
        """one snapshot's budget profiles (turb [19][NZ], vort [3][NZ], ke [30][NZ]),
        also folded into the handle's time averages"""
        NZ = self.dims.NZ
        turb, vort, ke = np.empty((19, NZ)), np.empty((3, NZ)), np.empty((30, NZ))
        self._check(self.lib.dnsg_budget2d_accumulate(self.h, _ptr(turb), _ptr(vort), _ptr(ke)))
        return turb, vort, ke

    def budget2d_average(self):
        """time averages of the folded snapshots (rms rows as mean squares) and their count"""
        NZ = self.dims.NZ
        turb, vort, ke = np.empty((19, NZ)), np.empty((3, NZ)), np.empty((30, NZ))
        cnt = ctypes.c_long()
        self._check(self.lib.dnsg_budget2d_average(self.h, _ptr(turb), _ptr(vort), _ptr(ke), ctypes.byref(cnt)))
        return turb, vort, ke, cnt.value

    def budget2d_reset(self):
        self._check(self.lib.dnsg_budget2d_reset(self.h))

    def run_timed(self, n, sections=None):
        """n steps without per-step host synchronisation: (device ms, diag).  With
        sections="all" (or one name of SECTIONS) also the average ms per launch of the
        kernel groups (or of that one group) over those same steps, keys as
        kernel_times"""
        d = (Diag * n)()
        ms = ctypes.c_double()
        if sections is None:
            self._check(self.lib.dnsg_run_timed(self.h, n, d, ctypes.byref(ms)))
            return ms.value, diag_array(d, n)
        group = -1 if sections == "all" else SECTIONS.index(sections)
        sec = np.zeros(5)
        self._check(self.lib.dnsg_run_timed_sections(self.h, n, d, ctypes.byref(ms), _ptr(sec), group))
        return ms.value, diag_array(d, n), dict(zip(SECTIONS, sec))

    def kernel_times(self, n):
        """average ms per launch: reductions, coarse CS, sampling, band CS, interface"""
        ms = np.zeros(5)
        self._check(self.lib.dnsg_kernel_times(self.h, n, _ptr(ms)))
        return dict(zip(SECTIONS, ms))

    def band_launch_substeps(self):
        """band substeps one band launch covers (2: both gratio-2 substeps in one pass)"""
        n = ctypes.c_int()
        self._check(self.lib.dnsg_band_launch_substeps(self.h, ctypes.byref(n)))
        return n.value

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
    def spectra(self, grid):
        """1-D energy spectra of grid 0/1/2: (ex [3][N0][nz], ey [3][N1/2+1][nz])"""
        d = self.dims
        N0, N1 = (d.Xc, d.Yc) if grid == 0 else (d.Xf, d.Yf)
        nz = d.zdc if grid == 0 else d.zdf
        ex = np.empty((3, N0, nz))
        ey = np.empty((3, N1 // 2 + 1, nz))
        self._check(self.lib.dnsg_spectra(self.h, grid, _ptr(ex), _ptr(ey)))
        return ex, ey

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


class Queue:
    """Serving queue of steady slip-length solves (dnsg_queue_*): submit channels
    as they arrive, advance, collect finished solves."""

    def __init__(self, **opts):
        self.lib = load()
        self.q = None
        o = QueueOpts()
        self.lib.dnsg_queue_opts_default(ctypes.byref(o))
        for k, v in opts.items():
            if not hasattr(o, k):
                raise KeyError(f"unknown queue option {k}")
            setattr(o, k, type(getattr(o, k))(v))
        self.opts = o
        q = _P()
        rc = self.lib.dnsg_queue_create(ctypes.byref(o), ctypes.byref(q))
        self.q = q
        self._check(rc)

    def _check(self, rc):
        if rc:
            raise DnsgError(rc, self.lib.dnsg_queue_last_error(self.q).decode())

    def close(self):
        if self.q:
            self.lib.dnsg_queue_destroy(self.q)
            self.q = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        return self.lib.dnsg_queue_clock(self.q)

    def submit(self, params, rid, t_arrival=-1.0):
        self._check(self.lib.dnsg_queue_submit(self.q, ctypes.byref(params), rid, t_arrival))

    def advance(self, steps, cap=256):
        buf = (SlipResult * cap)()
        nd = ctypes.c_int()
        self._check(self.lib.dnsg_queue_advance(self.q, steps, buf, cap, ctypes.byref(nd)))
        return [buf[i].asdict() for i in range(nd.value)]

    def graph_launches(self):
        """check intervals run as a replayed HIP graph so far"""
        n = ctypes.c_long()
        self._check(self.lib.dnsg_queue_graph_launches(self.q, ctypes.byref(n)))
        return n.value

    def info(self):
        p, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        self._check(self.lib.dnsg_queue_info(self.q, ctypes.byref(p), ctypes.byref(a), ctypes.byref(b)))
        return dict(pending=p.value, active=a.value, resident_bytes=b.value)


def solve_bytes(params):
    return load().dnsg_solve_bytes(ctypes.byref(params))


def hbm_probe(nbytes=2 << 30, kind=0, reps=5):
    """achievable HBM bandwidth (GB/s): kind 0 copy, 1 read, 2 write, 3 the
    collide-stream's data movement without its arithmetic"""
    g = ctypes.c_double()
    rc = load().dnsg_hbm_probe(nbytes, kind, reps, ctypes.byref(g))
    if rc:
        raise DnsgError(rc, "hbm probe")
    return g.value


def diag_array(d, n):
    return {f: np.array([getattr(d[i], f) for i in range(n)]) for f in DIAG_FIELDS}


def slip_request(xdv, ydv, zdv, gratio, pattern, period, gas_fraction, re_b, u0=0.05, precision="fp32", seed=0):
    """Parameters of one steady slip-length solve.

    Channel of coarse size xdv (streamwise) x ydv (spanwise) x zdv (wall-normal,
    including the two overlap layers); gratio 2 refines four coarse cells at each
    wall (zds = 8), gratio 1 keeps the near-wall band at the coarse resolution
    (zds = 4).  The wall pattern (stripes "rslip" or square posts "pslip") has
    period `period` fine cells and the given gas fraction.  The body force G gives
    the no-slip bulk velocity u0, and nu follows from Re_b = u0 H / nu with H the
    wall-to-wall height (the reference's convention, main:447).  The start is the
    no-slip laminar profile."""
    GR = int(gratio)
    zds = 8 if GR == 2 else max(4, GR + 3)
    zdf = zds + 1
    NZs = 2 * zdf + GR * (zdv - 4) + GR - 1
    H = (NZs - 2.0) / GR
    nu = u0 * H / re_b
    tau_c = 3.0 * nu + 0.5
    G = 12.0 * nu * u0 / (H * H)
    kw = dict(xdv=xdv, ydv=ydv, zdv=zdv, gratio=GR, zds=zds, tau_c=tau_c, ub=u0, force_mode="const", g_const=G,
              precision=precision, init_kind="laminar_noise", seed=seed, u_max=1.5 * u0, noise=0.0, CFL=1.0)
    P = int(period)
    if pattern == "rslip":
        g = min(P - 1, max(1, int(round(gas_fraction * P))))
        kw.update(pattern="rslip", gyplussize=g, wyplussize=P - g)
    elif pattern == "pslip":
        w = min(P - 1, max(1, int(round(P * (1.0 - gas_fraction) ** 0.5))))
        kw.update(pattern="pslip", gyplussize=P - w, wyplussize=w, gxplussize=P - w, wxplussize=w)
    else:
        kw.update(pattern="none")
    return make_params(**kw)


def slip_workload(n, seed, x_range=(32, 256), y_range=(32, 256), z_range=(16, 64), re_range=(10.0, 200.0),
                  precision="fp32"):
    """n slip-length requests drawn from one seed: coarse streamwise and spanwise
    sizes (multiples of 16) and wall-normal size uniform in their ranges, bands
    (gratio 2) on a random half, stripes or posts, gas fraction in [0.2, 0.9],
    period 8..64 fine cells dividing the patterned extents, Re_b uniform."""
    import random
    rng = random.Random(seed)
    out = []
    for i in range(n):
        xdv = rng.randrange(x_range[0], x_range[1] + 1, 16)
        ydv = rng.randrange(y_range[0], y_range[1] + 1, 16)
        zdv = rng.randint(*z_range)
        GR = 2 if rng.random() < 0.5 else 1
        pattern = "rslip" if rng.random() < 0.5 else "pslip"
        periods = [P for P in (8, 16, 32, 64) if (GR * ydv) % P == 0 and (pattern == "rslip" or (GR * xdv) % P == 0)]
        P = rng.choice(periods)
        phi = rng.uniform(0.2, 0.9)
        re_b = rng.uniform(*re_range)
        desc = dict(id=i, xdv=xdv, ydv=ydv, zdv=zdv, gratio=GR, pattern=pattern, period=P, gas_fraction=phi, re_b=re_b)
        out.append((desc, slip_request(xdv, ydv, zdv, GR, pattern, P, phi, re_b, precision=precision, seed=seed + i)))
    return out
