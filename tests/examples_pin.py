"""Solver input for the reference's own plunge example (test infrastructure).

QNM_Extraction/Examples holds a waveform the reference solver produced
(hm2_a0.7_thi000_thf090.0_n.dat: rows every 0.1 M = Fr * dt with the shipped
parm.h:1-8 and main.c:46, N = 16000, M = 256, rho in [-50, 50], dt = 0.002)
for the a = 0.7 prograde equatorial plunge in a0.7_thi000_thf090.0_n.traj,
and the QNM amplitudes fitted to it (excitation_coefficients_a70_thinc000_hm2.dat).
The trajectory file is trimmed (Scripts/Trim_Trajectory.py): it starts 50 M
before the orbital-frequency peak and is sampled every 0.02 M, while body.cu
reads two rows per step at dt/2 (body.cu:112-113) from the start of the run,
and its source switches on over ~100 M (kern.cu:2673).

`solver_rows` rebuilds what the solver reads:
  * t >= 2440.06 (the file): cubic Hermite interpolation of r (with dr/dt)
    and phi (with dphi/dt), cubic splines of dr/dt, dphi/dt, E, Lz; after the
    file's end (the particle sits on the horizon, dr/dt = -6e-14) r, the
    velocities, E and Lz stay constant and phi advances at dphi/dt;
  * t < 2440.06 (the approach the trimmed file lacks): the Ori-Thorne
    transition shape r = r_isco + R0 X(k (t - t_c)), X the universal solution
    of X'' = -X^2 - L with X -> sqrt(-L) (the function GenOT/OT_X_vs_L.dat
    tabulates), R0, k, t_c fitted to the file's first 30 M; E, Lz linear in t
    with the file's initial slope; dphi/dt the geodesic value at (r, E, Lz),
    which is what the file holds (to 1e-13); joined to the file with r and
    dr/dt continuous at 2440.06.

The solver's own clock starts at 0 at trajectory time `t0`; its output row
at step k corresponds to the reference's waveform time t0 + k dt - Xmax
(body.cu:207: timer - Xmax).
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
A = 0.7
T_FILE0 = 2440.06

_FIX = None


def fixture():
    global _FIX
    if _FIX is None:
        _FIX = dict(np.load(os.path.join(HERE, "golden", "examples_a07_traj.npz")))
    return _FIX


def isco_prograde(a):
    """Bardeen-Press-Teukolsky prograde equatorial ISCO radius (M = 1)."""
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = math.sqrt(3 * a * a + z1 * z1)
    return 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))


def ot_function():
    """X(L) on L in [-40, 3.4]: the solution of X'' = -X^2 - L that follows
    sqrt(-L) + 1/(8 L^2) as L -> -inf.  Returns a callable giving (X, X')."""
    from scipy.integrate import solve_ivp
    l0 = -40.0
    y0 = [math.sqrt(-l0) + 1 / (8 * l0 * l0), -0.5 / math.sqrt(-l0) - 1 / (4 * l0 ** 3)]
    sol = solve_ivp(lambda l, y: [y[1], -y[0] ** 2 - l], [l0, 3.4], y0, rtol=1e-12, atol=1e-13,
                    dense_output=True)
    return sol.sol


def geodesic_phidot(r, E, L, a=A):
    """dphi/dt of an equatorial Kerr geodesic with energy E, axial angular
    momentum L at radius r (M = 1)."""
    D = r * r - 2 * r + a * a
    s = r * r + a * a
    dtdl = (s * s / D - a * a) * E + a * L * (1 - s / D)
    dphdl = L - a * E + a / D * (E * s - a * L)
    return dphdl / dtdl


def _smoothstep(x):
    x = np.clip(x, 0.0, 1.0)
    return x * x * x * (10 - 15 * x + 6 * x * x), 30 * x * x * (1 - x) ** 2


def approach_fit(fit_span=30.0):
    """(R0, k, t_c) of r = r_isco + R0 X(k (t - t_c)) fitted to the file's
    first `fit_span` M."""
    from scipy.optimize import least_squares
    f = fixture()
    X = ot_function()
    rl = isco_prograde(A)
    sel = f["t"] < f["t"][0] + fit_span
    t, r = f["t"][sel], f["r"][sel]
    res = least_squares(lambda p: rl + p[0] * X(p[1] * (t - p[2]))[0] - r, [0.05, 0.05, 2430.0],
                        x_scale=[0.01, 0.01, 10.0])
    return tuple(res.x)


def solver_rows(t0, t_end, dt_out, join=20.0, approach="ot"):
    """(n, 10) .traj rows at t0 + j dt_out up to t_end.  approach "ot": the
    Ori-Thorne shape; "frozen": r continued backwards from the file's first
    row at its constant dr/dt, E and Lz held at their first values (a very
    different approach, for the sensitivity check)."""
    from scipy.interpolate import CubicHermiteSpline, CubicSpline
    f = fixture()
    tf = f["t"]
    t = t0 + dt_out * np.arange(int(round((t_end - t0) / dt_out)) + 1)
    rows = np.zeros((t.size, 10))
    rows[:, 0] = t
    rows[:, 2] = f["theta"]
    rows[:, 4] = f["thetadot"]
    rows[:, 9] = f["Q"]
    inside = (t >= tf[0]) & (t <= tf[-1])
    ti = t[inside]
    rows[inside, 1] = CubicHermiteSpline(tf, f["r"], f["rdot"])(ti)
    rows[inside, 6] = CubicHermiteSpline(tf, f["phi"], f["phidot"])(ti)
    for col, key in ((3, "rdot"), (5, "phidot"), (7, "E"), (8, "Lz")):
        rows[inside, col] = CubicSpline(tf, f[key])(ti)
    after = t > tf[-1]
    for col, key in ((1, "r"), (3, "rdot"), (5, "phidot"), (7, "E"), (8, "Lz")):
        rows[after, col] = f[key][-1]
    rows[after, 6] = f["phi"][-1] + f["phidot"][-1] * (t[after] - tf[-1])
    before = t < tf[0]
    if before.any():
        tb = t[before]
        r1, v1 = f["r"][0], f["rdot"][0]
        dEdt = (f["E"][1] - f["E"][0]) / (tf[1] - tf[0])
        dLdt = (f["Lz"][1] - f["Lz"][0]) / (tf[1] - tf[0])
        if approach == "ot":
            R0, k, tc = approach_fit()
            X = ot_function()
            xm = X(k * (tb - tc))
            rm, vm = isco_prograde(A) + R0 * xm[0], R0 * k * xm[1]
            x1 = X(k * (tf[0] - tc))
            dr, dv = r1 - (isco_prograde(A) + R0 * x1[0]), v1 - R0 * k * x1[1]
            w, dw = _smoothstep((tb - (tf[0] - join)) / join)
            off = dr + dv * (tb - tf[0])
            r = rm + w * off
            v = vm + w * dv + dw / join * off
            E = f["E"][0] + dEdt * (tb - tf[0])
            L = f["Lz"][0] + dLdt * (tb - tf[0])
        elif approach == "frozen":
            r = r1 + v1 * (tb - tf[0])
            v = np.full_like(tb, v1)
            E = np.full_like(tb, f["E"][0])
            L = np.full_like(tb, f["Lz"][0])
        else:
            raise ValueError(approach)
        phd = geodesic_phidot(r, E, L)
        # phi by the trapezoid rule backwards from the file's first phi
        tt = np.append(tb, tf[0])
        pd = np.append(phd, f["phidot"][0])
        seg = 0.5 * (pd[1:] + pd[:-1]) * np.diff(tt)
        phi = f["phi"][0] - np.cumsum(seg[::-1])[::-1]
        rows[before, 1], rows[before, 3], rows[before, 5], rows[before, 6] = r, v, phd, phi
        rows[before, 7], rows[before, 8] = E, L
    return rows
