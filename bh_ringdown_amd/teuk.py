"""Python view of libteuk's C ABI (include/teuk.h), for tests and benchmarks.

Mirrors the reference driver's interface (TD_GPU_code/HighOrder):
``grid`` is main.c:79-198 grid(), ``body`` is body.cu:25-225 body(), and
``Solver`` exposes the per-step operations body() is built from.  All
arithmetic runs in libteuk (HIP kernels on the GPU plus its host set-up);
this module only marshals arrays.  If the library has not been built, every
entry point raises -- there is no fallback path.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# TEUK_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("TEUK_LIB") or os.path.join(PKG, "lib", "libteuk.so")
PI15 = 3.14159265358979  # parm.h:12

_dp = ctypes.POINTER(ctypes.c_double)

TEUK_OK, TEUK_EINVAL, TEUK_EHIP, TEUK_ENOMEM, TEUK_EDIVERGED, TEUK_EIO, TEUK_ERANGE = 0, -1, -2, -3, -4, -5, -6
TEUK_ETOOBIG = -7
NOALLOC_BYTES = 64 << 20


class TeukError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {strerror(code)} ({code})")
        self.code = code


class Params(ctypes.Structure):
    """teuk_params (include/teuk.h); names follow parm.h:1-11 / main.c:55-61."""
    _fields_ = [
        ("N", ctypes.c_int), ("M", ctypes.c_int),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("mass", ctypes.c_double), ("a", ctypes.c_double),
        ("s", ctypes.c_double), ("m", ctypes.c_double),
        ("dx", ctypes.c_double), ("dtheta", ctypes.c_double), ("dt", ctypes.c_double),
        ("T", ctypes.c_double),
        ("Fr", ctypes.c_int), ("ptsx", ctypes.c_int), ("ptsy", ctypes.c_int),
        ("integrator", ctypes.c_int), ("j_extract", ctypes.c_int),
    ]

    def nsteps(self):
        return int(self.T / self.dt)

    def __repr__(self):
        return "Params(" + ", ".join(f"{n}={getattr(self, n)!r}" for n, _ in self._fields_) + ")"


class SourceStep(ctypes.Structure):
    """teuk_source_step: the per-step scalars body.cu:122-124 hands to sourced."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "timer", "start", "rp", "tp", "phip", "E", "lz", "Q",
        "drdt", "d2rdt2", "dthdt", "d2thdt2", "dphidt", "d2phidt2")]


class OutputRow(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("write", ctypes.c_int), ("t", ctypes.c_double),
                ("dedt", ctypes.c_double), ("h", ctypes.c_double * 14)]


# The same record as a numpy structured type (a zero-copy view of the rows).
OUTPUT_ROW_DTYPE = np.dtype([("k", np.int32), ("write", np.int32), ("t", np.float64), ("dedt", np.float64),
                             ("h", np.float64, (14,))])
assert OUTPUT_ROW_DTYPE.itemsize == ctypes.sizeof(OutputRow)


EXPORTS = [
    "teuk_grid", "teuk_default_params", "teuk_create", "teuk_destroy", "teuk_set_fields",
    "teuk_get_fields", "teuk_set_source", "teuk_set_trajectory", "teuk_step", "teuk_sync",
    "teuk_progress", "teuk_read_output", "teuk_read_psi", "teuk_enable_timing", "teuk_timing",
    "teuk_footprint", "teuk_body", "teuk_strerror", "teuk_trajectory_schedule", "teuk_traj_plunge",
    "teuk_set_tiling", "teuk_remaining", "teuk_batch_create", "teuk_batch_destroy", "teuk_batch_add",
    "teuk_batch_remove", "teuk_batch_set_tiling", "teuk_batch_step", "teuk_batch_sync", "teuk_batch_info",
    "teuk_source_coeffs", "teuk_source_patch", "teuk_grid_half", "teuk_create_staggered",
    "teuk_trajectory_schedule_staggered", "teuk_status", "teuk_solve_noalloc", "teuk_batch_set_cache_groups",
    "teuk_ot_transition", "teuk_set_theta_rows", "teuk_batch_set_theta_rows",
]

_lib = None


def lib():
    """Load libteuk.so (built by bh_ringdown_amd.build); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libteuk not built: {LIB_PATH} missing "
                           "(run python -c 'import __graft_entry__ as g; g.build()')")
    l = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(Params)
    vp = ctypes.c_void_p
    sig = {
        "teuk_grid": ([P, _dp, _dp, _dp], ctypes.c_int),
        "teuk_default_params": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double],
                                ctypes.c_int),
        "teuk_create": ([P, _dp, _dp, ctypes.POINTER(vp)], ctypes.c_int),
        "teuk_create_staggered": ([P, _dp, _dp, _dp, ctypes.POINTER(vp)], ctypes.c_int),
        "teuk_grid_half": ([P, _dp, _dp], ctypes.c_int),
        "teuk_trajectory_schedule_staggered": ([_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.POINTER(SourceStep)], ctypes.c_int),
        "teuk_destroy": ([vp], None),
        "teuk_set_fields": ([vp, _dp], ctypes.c_int),
        "teuk_get_fields": ([vp, _dp], ctypes.c_int),
        "teuk_set_source": ([vp, ctypes.POINTER(SourceStep), ctypes.c_int], ctypes.c_int),
        "teuk_set_trajectory": ([vp, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "teuk_step": ([vp, ctypes.c_int], ctypes.c_int),
        "teuk_sync": ([vp], ctypes.c_int),
        "teuk_status": ([vp], ctypes.c_int),
        "teuk_solve_noalloc": ([P, _dp, _dp, _dp, ctypes.POINTER(SourceStep), ctypes.c_int, _dp,
                                ctypes.POINTER(OutputRow), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                               ctypes.c_int),
        "teuk_progress": ([vp, ctypes.POINTER(ctypes.c_int), _dp], ctypes.c_int),
        "teuk_read_output": ([vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OutputRow),
                              ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "teuk_read_psi": ([vp, ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "teuk_enable_timing": ([vp, ctypes.c_int], ctypes.c_int),
        "teuk_timing": ([vp, ctypes.POINTER(ctypes.c_longlong), _dp], ctypes.c_int),
        "teuk_footprint": ([P, ctypes.c_int], ctypes.c_longlong),
        "teuk_body": ([P, _dp, _dp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int],
                      ctypes.c_int),
        "teuk_strerror": ([ctypes.c_int], ctypes.c_char_p),
        "teuk_trajectory_schedule": ([_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(SourceStep)], ctypes.c_int),
        "teuk_traj_plunge": ([ctypes.c_double] * 6 + [ctypes.c_ulonglong, _dp, ctypes.c_int,
                                                      ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "teuk_set_tiling": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "teuk_set_theta_rows": ([vp, ctypes.c_int], ctypes.c_int),
        "teuk_batch_set_theta_rows": ([vp, ctypes.c_int], ctypes.c_int),
        "teuk_ot_transition": ([ctypes.c_double, _dp, _dp], ctypes.c_int),
        "teuk_remaining": ([vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "teuk_source_coeffs": ([vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
        "teuk_source_patch": ([vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "teuk_batch_create": ([ctypes.c_int, ctypes.c_longlong, ctypes.POINTER(vp)], ctypes.c_int),
        "teuk_batch_destroy": ([vp], None),
        "teuk_batch_add": ([vp, vp], ctypes.c_int),
        "teuk_batch_remove": ([vp, vp], ctypes.c_int),
        "teuk_batch_set_tiling": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "teuk_batch_step": ([vp, ctypes.c_int], ctypes.c_int),
        "teuk_batch_set_cache_groups": ([vp, ctypes.c_longlong], ctypes.c_int),
        "teuk_batch_sync": ([vp], ctypes.c_int),
        "teuk_batch_info": ([vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(l, name)
        fn.argtypes = args
        fn.restype = res
    _lib = l
    return l


def strerror(code):
    return lib().teuk_strerror(int(code)).decode()


def _check(rc, what):
    if rc != TEUK_OK:
        raise TeukError(rc, what)


def _ptr(a):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


def default_params(N, M, xmin, xmax, a, m, dt_over_dx=0.5, T=0.0, **kw):
    p = Params()
    _check(lib().teuk_default_params(ctypes.byref(p), N, M, xmin, xmax, a, float(m), dt_over_dx, T),
           "teuk_default_params")
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def grid(p):
    """main.c:79-198 grid(): (x_c[N], r_c[N], theta[M])."""
    x = np.zeros(p.N); r = np.zeros(p.N); th = np.zeros(p.M)
    _check(lib().teuk_grid(ctypes.byref(p), _ptr(x), _ptr(r), _ptr(th)), "teuk_grid")
    return x, r, th


def grid_half(p):
    """Half points of the staggered scheme (integrator 2): (x_h[N], r_h[N])."""
    x = np.zeros(p.N); r = np.zeros(p.N)
    _check(lib().teuk_grid_half(ctypes.byref(p), _ptr(x), _ptr(r)), "teuk_grid_half")
    return x, r


def trajectory_schedule_staggered(rows, nsteps, dt, start=0.0):
    """The staggered scheme's source scalars, two per step (Old/body.cu:113-180)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = (SourceStep * (2 * nsteps))()
    _check(lib().teuk_trajectory_schedule_staggered(_ptr(rows), rows.shape[0], nsteps, start, dt, out),
           "teuk_trajectory_schedule_staggered")
    return list(out)


def trajectory_schedule(rows, nsteps, dt, start=0.0):
    """The per-step source scalars body.cu:97-129 derives from .traj rows."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = (SourceStep * nsteps)()
    _check(lib().teuk_trajectory_schedule(_ptr(rows), rows.shape[0], nsteps, start, dt, out),
           "teuk_trajectory_schedule")
    return list(out)


def plunge_trajectory(a, eta, iota, dt_out, t_final, seed=0, r0_over_isco=1.02):
    """Seeded transition-plus-plunge worldline as (n, 10) .traj rows."""
    cap = int(t_final / dt_out) + 4
    rows = np.zeros((cap, 10))
    n = ctypes.c_int()
    _check(lib().teuk_traj_plunge(a, eta, iota, r0_over_isco, dt_out, t_final, seed, _ptr(rows), cap,
                                  ctypes.byref(n)), "teuk_traj_plunge")
    return rows[: n.value].copy()


def ot_transition(L):
    """The universal Ori-Thorne transition function X(L) and X'(L)."""
    x = ctypes.c_double(); dx = ctypes.c_double()
    _check(lib().teuk_ot_transition(float(L), ctypes.byref(x), ctypes.byref(dx)), "teuk_ot_transition")
    return x.value, dx.value


def write_traj(path, rows):
    """Write rows in the .traj text format (numpy's %.18e, as the reference files)."""
    np.savetxt(path, rows, fmt="%.18e")


def footprint(p, nsteps=None):
    return int(lib().teuk_footprint(ctypes.byref(p), p.nsteps() if nsteps is None else nsteps))


class Solver:
    """One m-mode solve resident on the GPU (a libteuk handle)."""

    def __init__(self, p, r_c=None, theta=None, r_h=None):
        self.p = p
        if r_c is None or theta is None:
            _, r_c, theta = grid(p)
        self.r_c, self.theta = np.ascontiguousarray(r_c), np.ascontiguousarray(theta)
        h = ctypes.c_void_p()
        if r_h is not None:   # staggered scheme with given half-point radii
            self.r_h = np.ascontiguousarray(r_h, dtype=np.float64)
            _check(lib().teuk_create_staggered(ctypes.byref(p), _ptr(self.r_h), _ptr(self.r_c), _ptr(self.theta),
                                               ctypes.byref(h)), "teuk_create_staggered")
        else:
            _check(lib().teuk_create(ctypes.byref(p), _ptr(self.r_c), _ptr(self.theta), ctypes.byref(h)),
                   "teuk_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().teuk_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (4, self.p.M, self.p.N)

    def set_fields(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != self.shape:
            raise ValueError(f"fields must have shape {self.shape}")
        _check(lib().teuk_set_fields(self._h, _ptr(u)), "teuk_set_fields")

    def get_fields(self):
        u = np.zeros(self.shape)
        _check(lib().teuk_get_fields(self._h, _ptr(u)), "teuk_get_fields")
        return u

    def set_source(self, steps):
        if steps is None or len(steps) == 0:
            _check(lib().teuk_set_source(self._h, None, 0), "teuk_set_source")
            return
        arr = (SourceStep * len(steps))(*steps)
        _check(lib().teuk_set_source(self._h, arr, len(steps)), "teuk_set_source")

    def set_trajectory(self, rows, nsteps=None):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        if rows.ndim != 2 or rows.shape[1] != 10:
            raise ValueError("trajectory rows must be (n, 10)")
        n = self.p.nsteps() if nsteps is None else nsteps
        _check(lib().teuk_set_trajectory(self._h, _ptr(rows), rows.shape[0], n), "teuk_set_trajectory")

    def step(self, n):
        _check(lib().teuk_step(self._h, int(n)), "teuk_step")

    def sync(self):
        _check(lib().teuk_sync(self._h), "teuk_sync")

    def status(self):
        """TEUK_OK, or TEUK_EDIVERGED once the solve has failed (teuk_status)."""
        rc = lib().teuk_status(self._h)
        if rc not in (TEUK_OK, TEUK_EDIVERGED):
            _check(rc, "teuk_status")
        return rc

    def progress(self):
        k = ctypes.c_int(); t = ctypes.c_double()
        _check(lib().teuk_progress(self._h, ctypes.byref(k), ctypes.byref(t)), "teuk_progress")
        return k.value, t.value

    def _records_cap(self, first, max_rows):
        # records exist for the steps enqueued so far (one every Fr steps):
        # size the host buffers by those, not by the whole run
        done = self.p.nsteps() - self.remaining()
        return max(1, min(max_rows, (done + self.p.Fr - 1) // self.p.Fr - first + 1))

    def output(self, first=0, max_rows=1 << 20):
        n_have = ctypes.c_int()
        cap = self._records_cap(first, max_rows)
        rows = (OutputRow * cap)()
        _check(lib().teuk_read_output(self._h, first, cap, rows, ctypes.byref(n_have)), "teuk_read_output")
        v = np.frombuffer(rows, dtype=OUTPUT_ROW_DTYPE, count=n_have.value)
        return {"k": v["k"].astype(np.int64), "write": v["write"].astype(np.int64), "t": v["t"].copy(),
                "dedt": v["dedt"].copy(), "h": v["h"].copy().reshape(-1, 14)}

    def psi(self, first=0, max_rows=1 << 20):
        cap = self._records_cap(first, max_rows)
        buf = np.empty((cap, 2, self.p.M))
        n = ctypes.c_int()
        _check(lib().teuk_read_psi(self._h, first, cap, _ptr(buf), ctypes.byref(n)), "teuk_read_psi")
        return buf[: n.value].copy()

    def set_tiling(self, waves_per_block, waves_per_eu, theta_rows=None):
        """Stage launch shape; theta_rows (2, 4, 8) per wavefront if given."""
        _check(lib().teuk_set_tiling(self._h, int(waves_per_block), int(waves_per_eu)), "teuk_set_tiling")
        if theta_rows is not None:
            _check(lib().teuk_set_theta_rows(self._h, int(theta_rows)), "teuk_set_theta_rows")

    def source_coeffs(self, first=0, n=None):
        """Device-computed per-step source coefficients, shape (n, 12)."""
        if n is None:
            n = self.p.nsteps() - first
        c = np.zeros((n, 12))
        _check(lib().teuk_source_coeffs(self._h, int(first), int(n), _ptr(c)), "teuk_source_coeffs")
        return c

    def source_patch(self, k):
        """Device-computed source patch of step k: (patch[2, ptsy, ptsx], (row0, col0))."""
        buf = np.zeros((2, self.p.ptsy, self.p.ptsx))
        org = (ctypes.c_int * 2)()
        _check(lib().teuk_source_patch(self._h, int(k), _ptr(buf), org), "teuk_source_patch")
        return buf, (org[0], org[1])

    def remaining(self):
        n = ctypes.c_int()
        _check(lib().teuk_remaining(self._h, ctypes.byref(n)), "teuk_remaining")
        return n.value

    def enable_timing(self, on=True):
        _check(lib().teuk_enable_timing(self._h, 1 if on else 0), "teuk_enable_timing")

    def timing(self):
        n = ctypes.c_longlong(); ms = ctypes.c_double()
        _check(lib().teuk_timing(self._h, ctypes.byref(n), ctypes.byref(ms)), "teuk_timing")
        return n.value, ms.value


class Batch:
    """Solvers stepped together by the batched kernels (a libteuk batch).
    Holds references to its members; closing the batch leaves them valid."""

    def __init__(self, integrator=1, hbm_budget=0):
        h = ctypes.c_void_p()
        _check(lib().teuk_batch_create(int(integrator), int(hbm_budget), ctypes.byref(h)), "teuk_batch_create")
        self._h = h
        self.members = []

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().teuk_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, solver):
        """Admit a solver; returns False when it would exceed the HBM budget."""
        rc = lib().teuk_batch_add(self._h, solver._h)
        if rc == TEUK_ENOMEM:
            return False
        _check(rc, "teuk_batch_add")
        self.members.append(solver)
        return True

    def remove(self, solver):
        _check(lib().teuk_batch_remove(self._h, solver._h), "teuk_batch_remove")
        self.members.remove(solver)

    def set_tiling(self, waves_per_block=4, waves_per_eu=2, order=0, chunk=32, theta_rows=None):
        _check(lib().teuk_batch_set_tiling(self._h, int(waves_per_block), int(waves_per_eu), int(order),
                                           int(chunk)), "teuk_batch_set_tiling")
        if theta_rows is not None:
            _check(lib().teuk_batch_set_theta_rows(self._h, int(theta_rows)), "teuk_batch_set_theta_rows")

    def set_cache_groups(self, group_bytes):
        """Step members in groups whose field state fits group_bytes (0 = one group)."""
        _check(lib().teuk_batch_set_cache_groups(self._h, int(group_bytes)), "teuk_batch_set_cache_groups")

    def step(self, n):
        _check(lib().teuk_batch_step(self._h, int(n)), "teuk_batch_step")

    def sync(self):
        _check(lib().teuk_batch_sync(self._h), "teuk_batch_sync")

    def info(self):
        m = ctypes.c_int(); a = ctypes.c_int(); b = ctypes.c_longlong()
        _check(lib().teuk_batch_info(self._h, ctypes.byref(m), ctypes.byref(a), ctypes.byref(b)), "teuk_batch_info")
        return {"members": m.value, "active": a.value, "bytes": b.value}


def solve_noalloc(p, r_c, theta, nsteps, u0=None, steps=None, max_rows=None):
    """teuk_solve_noalloc: a whole solve in the library's fixed scratch memory.
    Returns (code, fields[4, M, N], records as Solver.output()); code is
    TEUK_OK, TEUK_EDIVERGED or TEUK_ETOOBIG (then fields and records are
    empty); other errors raise."""
    r_c = np.ascontiguousarray(r_c, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    if u0 is not None and np.shape(u0) != (4, p.M, p.N):
        raise ValueError("fields must have shape (4, M, N)")
    u0c = np.ascontiguousarray(u0, dtype=np.float64) if u0 is not None else None   # kept alive for the call
    u0p = _ptr(u0c) if u0c is not None else None
    arr = None
    if steps is not None and len(steps):
        if len(steps) != nsteps:
            raise ValueError("one source entry per step")
        arr = (SourceStep * len(steps))(*steps)
    cap = nsteps // p.Fr + 1 if max_rows is None else max_rows
    rows = (OutputRow * max(cap, 1))()
    u = np.zeros((4, p.M, p.N))
    n = ctypes.c_int()
    rc = lib().teuk_solve_noalloc(ctypes.byref(p), _ptr(r_c), _ptr(theta), u0p, arr, int(nsteps), _ptr(u), rows,
                                  cap, ctypes.byref(n))
    if rc not in (TEUK_OK, TEUK_EDIVERGED, TEUK_ETOOBIG):
        _check(rc, "teuk_solve_noalloc")
    k = n.value
    out = {"k": np.array([rows[i].k for i in range(k)], dtype=np.int64),
           "write": np.array([rows[i].write for i in range(k)], dtype=np.int64),
           "t": np.array([rows[i].t for i in range(k)]),
           "dedt": np.array([rows[i].dedt for i in range(k)]),
           "h": np.array([list(rows[i].h) for i in range(k)]).reshape(k, 14)}
    return rc, u, out


def body(p, r_c, theta, traj_path=None, out_path=None, full_path=None, verbose=False):
    """body.cu:25-225 body(): one full m-mode solve with the reference's file I/O."""
    enc = lambda s: s.encode() if s is not None else None
    _check(lib().teuk_body(ctypes.byref(p), _ptr(np.ascontiguousarray(r_c)), _ptr(np.ascontiguousarray(theta)),
                           enc(traj_path), enc(out_path), enc(full_path), 1 if verbose else 0), "teuk_body")
