"""The kept C driver (dnsg-run) end to end against the oracle's driver loop.

Both programs start from the same '-p' restart files (written by the oracle),
run the reference driver loop (dns-q-mgrd-all-patterns-xyc.c:1254-1375) and
write the reference's output files into their own directories.  Every file the
oracle writes must exist in the driver's directory too.

* constant body force: the lattice state (the ftemp half of lbf-c/slf/suf),
  the interface planes and force.NNNN are bit-equal; reductions (logs,
  statistics, velocity snapshots, the output_twoD2 profile) agree to 1e-10.
* flux-controlled force (the reference's own mode): every number agrees to
  1e-11 relative after 50 steps (the force comes from a sum taken in another
  order).
* '-s' rest start: the same comparison from the reference's rest state.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ORC_CLI, REPO
from helpers import PKG_DIR

DRIVER = os.path.join(PKG_DIR, "build", "dnsg-run")
UB = 0.0333333333333
MICRO = dict(xdv=16, ydv=8, zdv=12, gratio=2, zds=4, Rebs="20.0", ub=repr(UB), CFL="5.0", ITER=50,
             pattern=1, gyplussize=4, wyplussize=4)
SEEDED = dict(init_kind=1, seed=42, u_max="0.05", noise="0.05")
LOGS = ("flowrate.txt", "density.txt", "str-up.txt", "str-do.txt", "w-str.txt", "p-grad.txt")


def _args(kw):
    return [f"{k}={v}" for k, v in kw.items()]


def _run_pair(tmp_path, kw, mode):
    odir, gdir = tmp_path / "oracle", tmp_path / "driver"
    odir.mkdir()
    gdir.mkdir()
    subprocess.run([ORC_CLI, "init", str(odir)] + _args(kw), check=True)
    if mode == "p":
        for f in os.listdir(odir):
            shutil.copy(odir / f, gdir / f)
    subprocess.run([ORC_CLI, "run", str(odir)] + _args(kw) + ["threads=4"], check=True, capture_output=True)
    print(f"driver: dnsg-run -{mode} {' '.join(_args(kw))}", flush=True)
    r = subprocess.run([DRIVER, f"-{mode}"] + _args(kw) + ["results_dat=0"], cwd=gdir, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return odir, gdir, r.stdout


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300) if b.size else 0.0


def _compare(odir, gdir, exact_state, tol):
    names = sorted(os.listdir(odir))
    assert any(n.startswith("lbf-c.") for n in names)
    for name in names:
        assert (gdir / name).exists(), f"driver did not write {name}"
        if name.endswith(".txt"):
            a = np.loadtxt(gdir / name, ndmin=2)
            b = np.loadtxt(odir / name, ndmin=2)
            assert a.shape == b.shape, name
            assert np.array_equal(a[:, 0], b[:, 0]), name
            assert np.all(np.abs(a[:, 1] - b[:, 1]) <= 2e-12 + tol * np.abs(b[:, 1])), name
            continue
        a = np.fromfile(gdir / name, dtype=np.float64)
        b = np.fromfile(odir / name, dtype=np.float64)
        assert a.shape == b.shape, name
        if name.startswith(("lbf-c.", "lbf-slf.", "lbf-suf.")):
            # ftemp (the state) first; the second half is the reference's collision scratch
            a, b = a[: a.size // 2], b[: b.size // 2]
        state = name.startswith(("lbf-", "force."))
        if exact_state and state:
            assert np.array_equal(a, b), f"{name}: max diff {np.max(np.abs(a - b)):.3e}"
        elif name.startswith("force."):
            # G = ub - <rho u> cancels: bound it against ub, as the controller's input
            assert np.all(np.abs(a - b) <= tol * UB), f"{name}: {a} vs {b}"
        else:
            assert _rel(a, b) <= tol, f"{name}: rel {_rel(a, b):.3e}"


@pytest.mark.gpu
def test_driver_constant_force_bitwise(tmp_path, oracle_built):
    kw = dict(MICRO, **SEEDED, force_mode=1, g_const="2e-5")
    odir, gdir, out = _run_pair(tmp_path, kw, "p")
    _compare(odir, gdir, exact_state=True, tol=1e-10)
    assert "Average Density in the channel at end" in out


@pytest.mark.gpu
def test_driver_flux_mode(tmp_path, oracle_built):
    kw = dict(MICRO, **SEEDED)
    odir, gdir, _ = _run_pair(tmp_path, kw, "p")
    _compare(odir, gdir, exact_state=False, tol=1e-11)


@pytest.mark.gpu
def test_driver_filtered_injection(tmp_path, oracle_built):
    """FLTR (ftcdtransfer_lesfltr.c) through the driver: gratio 4, the thinnest bands"""
    kw = dict(MICRO, **SEEDED, gratio=4, zds=6, fltr=1, force_mode=1, g_const="2e-5", ITER=20)
    odir, gdir, _ = _run_pair(tmp_path, kw, "p")
    _compare(odir, gdir, exact_state=True, tol=1e-10)


@pytest.mark.gpu
def test_driver_uniform_start(tmp_path, oracle_built):
    """'-i': both programs interpolate the same uniform-grid file lbf.0000 and run on"""
    kw = dict(MICRO, force_mode=1, g_const="2e-5", ITER=10, xdu=24, ydu=12, zdu=13)
    odir, gdir = tmp_path / "oracle", tmp_path / "driver"
    odir.mkdir()
    gdir.mkdir()
    subprocess.run([ORC_CLI, "uniform", str(odir / "lbf.0000")] + _args(kw) + ["seed=77", "noise=0.1"], check=True)
    shutil.copy(odir / "lbf.0000", gdir / "lbf.0000")
    subprocess.run([ORC_CLI, "init", str(odir)] + _args(kw) + ["init_kind=2"], check=True)
    subprocess.run([ORC_CLI, "run", str(odir)] + _args(kw) + ["threads=4"], check=True, capture_output=True)
    print("driver: dnsg-run -i", flush=True)
    r = subprocess.run([DRIVER, "-i"] + _args(kw) + ["results_dat=0"], cwd=gdir, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    (odir / "lbf.0000").unlink()  # the input, not an output
    (gdir / "lbf.0000").unlink()
    _compare(odir, gdir, exact_state=True, tol=1e-10)


@pytest.mark.gpu
def test_driver_rest_start(tmp_path, oracle_built):
    kw = dict(MICRO, init_kind=0)
    odir, gdir, _ = _run_pair(tmp_path, kw, "s")
    _compare(odir, gdir, exact_state=False, tol=1e-11)


DEFINITIONS = """/* a run configuration in the format of the reference's definitions.h */
#define Rebs 20.0
#define ub 0.0333333333333
#define zdv  12
#define ydv  8
#define xdv  16
#define TSTR 0
#define TEND 1
#define CFL  5.0
#define gratio 2
#define zds 4
#define gyplussize 4
#define wyplussize 4
#define gxplussize 4
#define wxplussize 4
#define NOFLTR
//#define FLTR
//#define PSLIP
#define RSLIP
//#define SSLIP
#define XDIM 4
#define YDIM 2
#define ITER 50
#define STRIPE_COUNT "64"
#define Fs(a,i,j,k,xd,yd,zd) (*(a+(i*yd*zd)+(j*zd)+k))
"""


def test_driver_parses_definitions(tmp_path):
    """`#define` lines, the pattern switches and ITER are read; comments, macros with
    arguments and the MPI/IO settings are ignored; the derived sizes are the oracle's"""
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    (tmp_path / "definitions.h").write_text(DEFINITIONS)
    r = subprocess.run([DRIVER, "-s", str(tmp_path / "definitions.h"), "print_params=1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    got = dict(line.split("=", 1) for line in r.stdout.split())
    assert got["pattern"] == "1" and got["ITER"] == "50" and got["xdv"] == "16" and got["zds"] == "4"
    assert float(got["ub"]) == UB and float(got["CFL"]) == 5.0 and got["force_mode"] == "0"
    assert got["fltr"] == "0"
    from helpers import Oracle, dnsg
    o = Oracle(dnsg.make_params(**MICRO))
    assert int(got["NZs"]) == o.dims.NZs and int(got["NZ"]) == o.dims.NZ
    assert int(got["utime"]) == o.dims.utime and int(got["up"]) == o.dims.up
    assert float(got["tau_c"]) == o.dims.tau_c and float(got["tau_f"]) == o.dims.tau_f
    o.close()
    # the FLTR switch selects the filtered injection, which needs gratio 4
    fl = DEFINITIONS.replace("//#define FLTR", "#define FLTR").replace("#define NOFLTR", "//#define NOFLTR")
    (tmp_path / "definitions.h").write_text(fl.replace("#define gratio 2", "#define gratio 4")
                                            .replace("#define zds 4", "#define zds 6"))
    r = subprocess.run([DRIVER, "-s", str(tmp_path / "definitions.h"), "print_params=1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    assert dict(line.split("=", 1) for line in r.stdout.split())["fltr"] == "1"
    (tmp_path / "definitions.h").write_text(fl)
    r = subprocess.run([DRIVER, "-s", str(tmp_path / "definitions.h"), "print_params=1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=30)
    assert r.returncode != 0  # FLTR with gratio 2 is rejected


@pytest.mark.gpu
def test_driver_runs_from_definitions_file(tmp_path, oracle_built):
    """`dnsg-run -s definitions.h` (the reference's own invocation) against the oracle"""
    kw = dict(MICRO, init_kind=0)
    odir, gdir = tmp_path / "oracle", tmp_path / "driver"
    odir.mkdir()
    gdir.mkdir()
    subprocess.run([ORC_CLI, "init", str(odir)] + _args(kw), check=True)
    subprocess.run([ORC_CLI, "run", str(odir)] + _args(kw) + ["threads=4"], check=True, capture_output=True)
    (gdir / "definitions.h").write_text(DEFINITIONS)
    print("driver: dnsg-run -s definitions.h", flush=True)
    r = subprocess.run([DRIVER, "-s", "definitions.h", "results_dat=0"], cwd=gdir, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    (gdir / "definitions.h").unlink()
    _compare(odir, gdir, exact_state=False, tol=1e-11)


@pytest.mark.gpu
def test_driver_writes_spectra(tmp_path, oracle_built):
    """spectra=1: the E*-{c,fl,fu}.NNNN files of every velocity snapshot, equal to the
    spectra of the vel-* files written beside them (the reference tool's input)"""
    from test_gpu_spectra import reference_spectra
    kw = dict(MICRO, **SEEDED, force_mode=1, g_const="2e-5", ITER=3)
    odir, gdir, _ = _run_pair(tmp_path, kw, "p")
    r = subprocess.run([DRIVER, "-p"] + _args(kw) + ["results_dat=0", "spectra=1"], cwd=gdir, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    from helpers import dnsg
    d = dnsg.dims_of(dnsg.make_params(**MICRO))
    shapes = {"c": (d.Xc, d.Yc, d.zdc), "fl": (d.Xf, d.Yf, d.zdf), "fu": (d.Xf, d.Yf, d.zdf)}
    velname = {"c": "vel-c", "fl": "vel-slf", "fu": "vel-suf"}
    for idx in (0, 3):
        for tag, (X, Y, Z) in shapes.items():
            vel = np.fromfile(gdir / f"{velname[tag]}.{idx:04d}").reshape(X, Y, Z, 3)
            rx, ry = reference_spectra(vel)
            for c, name in enumerate(("uu", "vv", "ww")):
                ex = np.fromfile(gdir / f"E{name}x-{tag}.{idx:04d}").reshape(X, Z)
                ey = np.fromfile(gdir / f"E{name}y-{tag}.{idx:04d}").reshape(Y // 2 + 1, Z)
                assert np.max(np.abs(ex - rx[c])) <= 1e-12 * max(np.max(np.abs(rx)), 1e-300), (tag, name)
                assert np.max(np.abs(ey - ry[c])) <= 1e-12 * max(np.max(np.abs(ry)), 1e-300), (tag, name)


@pytest.mark.gpu
def test_driver_writes_budget(tmp_path, oracle_built):
    """budget=1: the TKE-budget tool's files of every snapshot, against the oracle's
    budget pass over the velocity/density files written beside them"""
    kw = dict(MICRO, **SEEDED, force_mode=1, g_const="2e-5", ITER=3)
    odir, gdir, _ = _run_pair(tmp_path, kw, "p")
    r = subprocess.run([DRIVER, "-p"] + _args(kw) + ["results_dat=0", "budget=1"], cwd=gdir, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    post = gdir / "post"
    post.mkdir()
    subprocess.run([ORC_CLI, "budget", str(post)] + _args(kw) + ["TSTR=0", "TEND=3"], check=True)
    from helpers import budget_scales, dnsg
    d_nz = dnsg.dims_of(dnsg.make_params(**MICRO)).NZ
    names = sorted(os.listdir(post))
    assert len(names) == 3 * 4 + 2
    for name in names:
        a = np.fromfile(gdir / name)
        b = np.fromfile(post / name)
        assert a.shape == b.shape, name
        kind = "plane" if name.startswith(("ufl", "vfl")) else {"tu": "turb", "vo": "vort", "KE": "ke"}[name[:2]]
        shape = (1, -1) if kind == "plane" else (-1, d_nz)
        a, b = a.reshape(shape), b.reshape(shape)
        scale = budget_scales(kind, b) + 1e-300
        assert np.all(np.max(np.abs(a - b), axis=1) <= 1e-10 * scale), name


POSTPROC = os.path.join(REPO, "oracle", "_ref", "postproc_micro", "postprocess-budget")


def _dat_numbers(path):
    out = []
    for line in open(path):
        try:
            out.extend(float(x) for x in line.split())
        except ValueError:
            continue  # the Tecplot title / variables / zone lines
    return np.array(out)


@pytest.mark.gpu
def test_reference_postprocessor_reads_driver_files(tmp_path, oracle_built):
    """The reference's own TKE-budget post-processor (TKE-budget/postprocessing, built
    from its sources by oracle/build_postproc_ref.sh; it averages the per-snapshot
    budget files over time, time_stat.c / vorticity.c / Ebudget.c, and writes them in
    wall units with u_tau from p-grad.txt and U_b from flowrate.txt) run unchanged in
    the driver's run directory, and in a directory holding the oracle's files of the
    same run: it reads both and writes the same 81 profile files, every number within
    1e-8 + 1e-6 relative (the per-snapshot inputs agree to 1e-10 of each profile's
    scale; the tool prints 9 decimals).  The density skewness and kurtosis are left
    out (skew-p*, kurt-p* and the last column of skew-all / kurt-all): the seeded start
    has density fluctuations at rounding level, so those moments are ratios of rounding
    noise and differ between any two correct implementations."""
    if not os.path.exists(POSTPROC):
        pytest.skip("reference post-processor not built (oracle/build_postproc_ref.sh needs the reference sources)")
    kw = dict(MICRO, **SEEDED, force_mode=1, g_const="2e-5", ITER=3)
    odir, gdir, _ = _run_pair(tmp_path, kw, "p")
    r = subprocess.run([DRIVER, "-p"] + _args(kw) + ["results_dat=0", "budget=1"], cwd=gdir, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    post = odir / "post"
    post.mkdir()
    subprocess.run([ORC_CLI, "budget", str(post)] + _args(kw) + ["TSTR=0", "TEND=3"], cwd=odir, check=True)
    for f in ("p-grad.txt", "flowrate.txt"):
        shutil.copy(odir / f, post / f)
    for d in (gdir, post):
        rp = subprocess.run([POSTPROC], cwd=d, capture_output=True, text=True, timeout=120)
        assert rp.returncode == 0, (str(d), rp.stdout[-1000:], rp.stderr[-1000:])
    names = sorted(n for n in os.listdir(post) if n.endswith(".dat"))
    assert len(names) == 81, len(names)
    for name in names:
        if name.startswith(("skew-p", "kurt-p")):
            continue
        a, b = _dat_numbers(gdir / name), _dat_numbers(post / name)
        assert a.shape == b.shape and a.size > 0, name
        if name in ("skew-all.dat", "kurt-all.dat"):  # rows: k, z+, u, w, v, p
            keep = np.arange(a.size) % 6 != 5
            a, b = a[keep], b[keep]
        assert np.array_equal(np.isfinite(a), np.isfinite(b)), name
        f = np.isfinite(b)
        assert np.all(np.abs(a[f] - b[f]) <= 1e-8 + 1e-6 * np.abs(b[f])), (name, np.max(np.abs(a[f] - b[f])))


def test_driver_rejects_unknown_key(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    r = subprocess.run([DRIVER, "-s", "no_such_key=1"], cwd=tmp_path, capture_output=True, text=True, timeout=30)
    assert r.returncode != 0
    assert "no_such_key" in (r.stdout + r.stderr)
