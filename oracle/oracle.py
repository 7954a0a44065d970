"""ctypes view of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/libteuk_oracle.so`` (parity build) or
``oracle/libteuk_oracle_omp.so`` (timing build) and exposes the restated
reference update loop of TD_GPU_code/HighOrder to the tests, the smoke check
and bench.py's cpu_baseline leg.  Nothing in the product imports this module.

Also loads ``oracle/_ref/libref_host.so`` -- the reference's own main.c and
proj.c compiled as they lie (parm.h fixes N=16000, M=256, rho in [-50,50]) --
when it has been built, to pin grid() and project() bitwise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int), ("M", ctypes.c_int),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("mass", ctypes.c_double), ("a", ctypes.c_double),
        ("s", ctypes.c_double), ("m", ctypes.c_double),
        ("dx", ctypes.c_double), ("dtheta", ctypes.c_double), ("dt", ctypes.c_double),
        ("integrator", ctypes.c_int), ("j_extract", ctypes.c_int),
        ("ptsx", ctypes.c_int), ("ptsy", ctypes.c_int),
        ("nthreads", ctypes.c_int),
    ]


class SrcStep(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "timer", "start", "rp", "tp", "phip", "E", "lz", "Q",
        "drdt", "d2rdt2", "dthdt", "d2thdt2", "dphidt", "d2phidt2")]


PI15 = 3.14159265358979  # parm.h:12


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run: make -C oracle)")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(Params)
    lib.orc_grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
    lib.orc_coeffs.argtypes = [P, _dp, _dp, _dp]
    lib.orc_rhs.argtypes = [P, _dp, _dp, _dp, _dp, _dp]
    lib.orc_bc_theta.argtypes = [P, _dp]
    lib.orc_bc_rho.argtypes = [P, _dp]
    lib.orc_step.argtypes = [P, _dp, _dp, _dp, _dp, _dp]
    lib.orc_source.argtypes = [P, _dp, _dp, ctypes.POINTER(SrcStep), _dp]
    lib.orc_fact.argtypes = [ctypes.c_double]
    lib.orc_fact.restype = ctypes.c_double
    lib.orc_swsh.argtypes = [ctypes.c_double] * 3
    lib.orc_swsh.restype = ctypes.c_double
    lib.orc_project.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.orc_project.restype = ctypes.c_double
    lib.orc_run.argtypes = [P, _dp, _dp, _dp, ctypes.c_int, ctypes.POINTER(SrcStep), _dp, _dp, _dp]
    lib.orc_run.restype = ctypes.c_int
    lib.orc_t4_coeffs.argtypes = [_dp, _dp]
    lib.orc_exp_neg.argtypes = [ctypes.c_double]
    lib.orc_exp_neg.restype = ctypes.c_double
    lib.orc_grid_half.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _dp, _dp]
    lib.orc_coeffs_so2.argtypes = [P, _dp, _dp, _dp, _dp, _dp]
    lib.orc_step_so2.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_run_so2.argtypes = [P, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.POINTER(SrcStep), _dp, _dp, _dp]
    lib.orc_run_so2.restype = ctypes.c_int
    return lib


_libs = {}


def lib(timing=False):
    name = "libteuk_oracle_omp.so" if timing else "libteuk_oracle.so"
    if name not in _libs:
        _libs[name] = _load(name)
    return _libs[name]


# Threads for the point loops (results do not depend on the count).
DEFAULT_THREADS = max(1, min(16, os.cpu_count() or 1))


def make_params(N, M, xmin, xmax, a, m, dt=None, mass=1.0, integrator=0, j_extract=-1,
                ptsx=128, ptsy=128, nthreads=DEFAULT_THREADS, dt_over_dx=0.5):
    dx = (xmax - xmin) / N
    return Params(N, M, xmin, xmax, mass, a, -2.0, float(m), dx, PI15 / M,
                  dt if dt is not None else dt_over_dx * dx,
                  integrator, j_extract, ptsx, ptsy, nthreads)


def grid(p):
    x = np.zeros(p.N); r = np.zeros(p.N); th = np.zeros(p.M)
    lib().orc_grid(p.N, p.M, p.xmin, p.xmax, p.mass, p.a, _ptr(x), _ptr(r), _ptr(th))
    return x, r, th


def coeffs(p, theta, r_c):
    c = np.zeros((9, p.M, p.N))
    lib().orc_coeffs(ctypes.byref(p), _ptr(theta), _ptr(r_c), _ptr(c))
    return c


def rhs(p, coef, theta, u, src=None):
    out = np.zeros_like(u)
    lib().orc_rhs(ctypes.byref(p), _ptr(coef), _ptr(theta), _ptr(u),
                  _ptr(src) if src is not None else None, _ptr(out))
    return out


def bc(p, u):
    lib().orc_bc_theta(ctypes.byref(p), _ptr(u))
    lib().orc_bc_rho(ctypes.byref(p), _ptr(u))
    return u


def step(p, coef, theta, u, src=None, work=None):
    if work is None:
        work = np.zeros((16, p.M, p.N))
    lib().orc_step(ctypes.byref(p), _ptr(coef), _ptr(theta), _ptr(u),
                   _ptr(src) if src is not None else None, _ptr(work))
    return u


def source(p, theta, r_c, st):
    out = np.zeros((2, p.M, p.N))
    lib().orc_source(ctypes.byref(p), _ptr(theta), _ptr(r_c), ctypes.byref(st), _ptr(out))
    return out


T4_IN = ["mass", "a", "E", "lz", "nmu", "mm", "rp", "drdt", "d2rdt2", "dthdt", "d2thdt2",
         "dphidt", "d2phidt2", "stp", "ctp", "cmphi", "smphi"]


def t4_coeffs(state):
    """Canonical T_4 coefficients x e^{-i m phi}/pi for a particle state dict
    (keys of T4_IN, with tp and phip instead of the trigonometric values)."""
    import math
    v = dict(state)
    v["stp"], v["ctp"] = math.sin(v["tp"]), math.cos(v["tp"])
    v["cmphi"], v["smphi"] = math.cos(v["mm"] * v["phip"]), math.sin(v["mm"] * v["phip"])
    inp = np.array([float(v[k]) for k in T4_IN])
    out = np.zeros(12)
    lib().orc_t4_coeffs(_ptr(inp), _ptr(out))
    return out


def swsh(l, m, th):
    return lib().orc_swsh(float(l), float(m), float(th))


def project(qq, M, l, m):
    buf = np.zeros(M + 1)
    buf[:M] = qq[:M]
    return lib().orc_project(_ptr(buf), M, float(l), float(m))


def run(p, r_c, theta, u, nsteps, src_steps=None, timing=False, want_h=True):
    """Advance u in place by nsteps; returns (status, psi_hist, h_hist, dedt_hist)."""
    M = p.M
    psi = np.zeros((nsteps, 2, M))
    h = np.zeros((nsteps, 14)) if want_h else None
    dedt = np.zeros(nsteps)
    steps = None
    if src_steps is not None:
        arr = (SrcStep * nsteps)(*src_steps)
        steps = ctypes.cast(arr, ctypes.POINTER(SrcStep))
    rc = lib(timing).orc_run(ctypes.byref(p), _ptr(r_c), _ptr(theta), _ptr(u), nsteps, steps,
                             _ptr(psi), _ptr(h) if h is not None else None, _ptr(dedt))
    return rc, psi, h, dedt


# ---------------------------------------------- staggered scheme (Old/) --
def grid_half(N, xmin, xmax, mass, a):
    """Half-point rho_h = Xmin + (l - 1/2) dx and radii r_h of the staggered scheme."""
    x = np.zeros(N); r = np.zeros(N)
    lib().orc_grid_half(N, xmin, xmax, mass, a, _ptr(x), _ptr(r))
    return x, r


def coeffs_so2(p, theta, r_h, r_c):
    ch = np.zeros((9, p.M, p.N)); cc = np.zeros((9, p.M, p.N))
    lib().orc_coeffs_so2(ctypes.byref(p), _ptr(theta), _ptr(r_h), _ptr(r_c), _ptr(ch), _ptr(cc))
    return ch, cc


def step_so2(p, coef_h, coef_c, theta, u, src_h=None, src_c=None, work=None):
    """One staggered step in place; pass the same work (12 planes, zero at
    first) from step to step."""
    if work is None:
        work = np.zeros((12, p.M, p.N))
    lib().orc_step_so2(ctypes.byref(p), _ptr(coef_h), _ptr(coef_c), _ptr(theta), _ptr(u),
                       _ptr(src_h) if src_h is not None else None,
                       _ptr(src_c) if src_c is not None else None, _ptr(work))
    return u


def run_so2(p, r_h, r_c, theta, u, nsteps, src_steps=None, want_h=True, timing=False):
    """Staggered-scheme run (two source entries per step); as run()."""
    M = p.M
    psi = np.zeros((nsteps, 2, M))
    h = np.zeros((nsteps, 14)) if want_h else None
    dedt = np.zeros(nsteps)
    steps = None
    if src_steps is not None:
        assert len(src_steps) >= 2 * nsteps
        arr = (SrcStep * len(src_steps))(*src_steps)
        steps = ctypes.cast(arr, ctypes.POINTER(SrcStep))
    rc = lib(timing).orc_run_so2(ctypes.byref(p), _ptr(r_h), _ptr(r_c), _ptr(theta), _ptr(u), nsteps, steps,
                           _ptr(psi), _ptr(h) if h is not None else None, _ptr(dedt))
    return rc, psi, h, dedt


# ---------------------------------------------------------------- reference --
REF_N, REF_M, REF_XMIN, REF_XMAX = 16000, 256, -50.0, 50.0  # parm.h:1-4


def ref_host():
    """The reference main.c/proj.c built as they lie, or None if not built."""
    path = os.path.join(HERE, "_ref", "libref_host.so")
    if not os.path.exists(path):
        return None
    if "ref" not in _libs:
        l = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        l.grid.argtypes = []
        l.grid.restype = None
        l.swsh.argtypes = [ctypes.c_double] * 3
        l.swsh.restype = ctypes.c_double
        l.project.argtypes = [_dp, ctypes.c_double, ctypes.c_double]
        l.project.restype = ctypes.c_double
        _libs["ref"] = l
    return _libs["ref"]


def ref_grid(mass, a):
    """main.c:79-198 grid() of the reference at its compiled size."""
    l = ref_host()
    ctypes.c_double.in_dll(l, "mass").value = mass
    ctypes.c_double.in_dll(l, "aa").value = a
    ctypes.c_double.in_dll(l, "dx").value = (REF_XMAX - REF_XMIN) / REF_N
    ctypes.c_double.in_dll(l, "dtheta").value = PI15 / REF_M
    l.grid()
    x = np.ctypeslib.as_array((ctypes.c_double * REF_N).in_dll(l, "x_c")).copy()
    r = np.ctypeslib.as_array((ctypes.c_double * REF_N).in_dll(l, "r_c")).copy()
    th = np.ctypeslib.as_array((ctypes.c_double * REF_M).in_dll(l, "theta")).copy()
    return x, r, th


def ref_project(qq, l, m):
    """proj.c:37-75 project() of the reference (M=256)."""
    buf = np.zeros(REF_M + 1)
    buf[:REF_M] = qq[:REF_M]
    return ref_host().project(_ptr(buf), float(l), float(m))
