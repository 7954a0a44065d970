"""Pin the oracle by what the mathematics fixes (no reference build needed).

Each check is chosen so that a plausible restatement mistake -- a wrong
stencil weight, a sign in the complex assembly, a boundary fill in the wrong
order, a wrong Runge-Kutta weight -- fails it.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

PI15 = O.PI15


def params(N=128, M=16, a=0.5, m=2, xmin=-100.0, xmax=300.0, integrator=0):
    return O.make_params(N, M, xmin, xmax, a, m, integrator=integrator)


def zero_coef(p):
    return np.zeros((9, p.M, p.N))


def theta_of(p):
    return np.arange(p.M) * (PI15 / p.M)


INNER = (slice(3, -3), slice(4, -4))  # rows 3..M-4, KO-active columns 4..N-5


# ------------------------------------------------------------------ grid ----
@pytest.mark.parametrize("a,xmin,xmax", [(0.5, -100.0, 300.0), (0.9, -400.0, 1200.0), (0.0, -50.0, 50.0)])
def test_grid_newton_residual(a, xmin, xmax):
    p = params(N=2048, a=a, xmin=xmin, xmax=xmax)
    x, r, th = O.grid(p)
    rp = 1 + math.sqrt(1 - a * a)
    rm = 1 - math.sqrt(1 - a * a)
    rin = xmin + 0.64 * (xmax - xmin)
    om = np.where(x < rin, 1.0, 1.0 - ((x - rin) / (xmax + 1e-6 - rin)) ** 4)
    target = x / om
    ok = r > rp * (1 + 1e-9)          # points resolved away from the horizon
    rr = r[ok]
    if a == 0.0:
        rstar = rr + 2 * np.log(rr / 2 - 1)
    else:
        rstar = rr + (rp * rp + a * a) / (rp - rm) * np.log(rr / rp - 1) - (rm * rm + a * a) / (rp - rm) * np.log(rr / rm - 1)
    # Newton stops at |dr|/r <= 1e-12; the r* residual is bounded by that step
    # times dr*/dr = (r^2+a^2)/Delta.
    jac = (rr * rr + a * a) / (rr * rr - 2 * rr + a * a)
    assert np.all(np.abs(rstar - target[ok]) <= 2e-12 * rr * jac + 1e-9)
    assert np.all(np.diff(r) > 0) or np.all(np.diff(r[ok]) > 0)
    # unresolved points sit at the horizon
    assert np.all(r[~ok] >= rp)


# ------------------------------------------------------------ projections ---
def test_swsh_closed_forms():
    for th in np.linspace(0.05, 3.0, 13):
        c = math.cos(th)
        y22 = math.sqrt(5 / (64 * PI15)) * (1 + c) ** 2
        y2m2 = math.sqrt(5 / (64 * PI15)) * (1 - c) ** 2
        assert abs(O.swsh(2, 2, th) - y22) <= 1e-14
        assert abs(O.swsh(2, -2, th) - y2m2) <= 1e-14
        # l < 2 has no spin-weight -2 harmonic
        assert O.swsh(1, 0, th) == 0.0


def test_swsh_orthonormal():
    xg, wg = np.polynomial.legendre.leggauss(200)
    th = np.arccos(xg)
    for m in (-2, 0, 1, 3):
        ls = list(range(max(2, abs(m)), 8))
        Y = np.array([[O.swsh(l, m, t) for t in th] for l in ls])
        G = 2 * math.pi * (Y * wg) @ Y.T
        assert np.allclose(G, np.eye(len(ls)), atol=1e-12)


@pytest.mark.parametrize("m", [-2, 2, 3])
def test_project_recovers_modes(m):
    """project() is 4 pi x Boole quadrature of qq Y sin: 2 delta_ll' for qq = Y_l'm."""
    M = 128
    th = np.arange(M) * (PI15 / M)
    for lp in range(max(2, abs(m)), abs(m) + 4):
        qq = np.array([O.swsh(lp, m, t) for t in th])
        for l in range(max(2, abs(m)), abs(m) + 4):
            got = O.project(qq, M, l, m)
            assert abs(got - (2.0 if l == lp else 0.0)) < 2e-5


# ------------------------------------------------------------------- rhs ----
def test_rho_derivative_exact_on_sextics_and_ko_kills_septics():
    p = params(N=128, M=16)
    coef = zero_coef(p)
    coef[0] = -1.0                               # bb = -1  ->  rhs_q = +q_x + p - KO
    x = p.xmin + np.arange(p.N) * p.dx
    s = (x - x.mean()) / 40.0
    f = 0.3 + s - 0.7 * s ** 2 + 0.2 * s ** 3 + 0.5 * s ** 4 - 0.1 * s ** 5 + 0.05 * s ** 6
    df = (1 - 1.4 * s + 0.6 * s ** 2 + 2.0 * s ** 3 - 0.5 * s ** 4 + 0.3 * s ** 5) / 40.0
    u = np.zeros((4, p.M, p.N))
    u[0] = f
    r = O.rhs(p, coef, theta_of(p), u)
    # literal 0.016666666666667 differs from 1/60 by 2e-14 relative
    assert np.allclose(r[0][INNER], np.broadcast_to(df, (p.M, p.N))[INNER], rtol=1e-11, atol=1e-14)
    g = f + 0.01 * s ** 7                         # KO annihilates degree <= 7
    u[0] = g
    coef[0] = 0.0
    r = O.rhs(p, coef, theta_of(p), u)
    assert np.max(np.abs(r[0][INNER])) < 1e-9


def test_ko_checkerboard_and_range():
    p = params(N=128, M=16)
    u = np.zeros((4, p.M, p.N))
    u[1] = (-1.0) ** np.arange(p.N)               # highest rho frequency in q_im
    r = O.rhs(p, zero_coef(p), theta_of(p), u)
    expect = -0.1 * 256.0 * 0.0625 / p.dt * (-1.0) ** np.arange(p.N)
    assert np.allclose(r[1][3:-3, 4:-4], expect[4:-4], rtol=1e-13)
    assert np.all(r[1][3:-3, 3] == 0.0) and np.all(r[1][3:-3, p.N - 4] == 0.0)


def test_theta_operator_on_polynomials():
    p = params(N=128, M=32)
    coef = zero_coef(p)
    coef[1] = 1.0                                 # cc = 1  ->  rhs_p = q_yy + cot q_y
    th = theta_of(p)
    t = th - 1.3
    f = 1 + 0.4 * t - 0.3 * t ** 2 + 0.2 * t ** 3 + 0.1 * t ** 4 - 0.05 * t ** 5 + 0.02 * t ** 6
    d1 = 0.4 - 0.6 * t + 0.6 * t ** 2 + 0.4 * t ** 3 - 0.25 * t ** 4 + 0.12 * t ** 5
    d2 = -0.6 + 1.2 * t + 1.2 * t ** 2 - 1.0 * t ** 3 + 0.6 * t ** 4
    u = np.zeros((4, p.M, p.N))
    u[0] = f[:, None]
    r = O.rhs(p, coef, th, u)
    cot = np.tan(0.5 * PI15 - th)
    expect = d2 + cot * d1
    assert np.allclose(r[2][3:-3, 3:-3], expect[3:-3, None], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("plane,sign_re,sign_im", [
    (3, "-p", "-p"),      # a_re: rhs_p = -(a_re) p
    (4, "+pi,-pr", None),  # a_im
    (7, "-q", "-q"),      # v_re
    (8, "+qi,-qr", None),  # v_im
])
def test_complex_assembly_signs(plane, sign_re, sign_im):
    """rhs_p = src + ee p_x + cc L q + d q_x - a p - v q with complex a, d, v."""
    p = params(N=128, M=16)
    rng = np.random.default_rng(plane)
    u = np.zeros((4, p.M, p.N))
    u[:, :, :] = rng.standard_normal((4, 1, 1))   # constant fields: no derivatives
    coef = zero_coef(p)
    coef[plane] = 1.0
    r = O.rhs(p, coef, theta_of(p), u)
    qr, qi, pr, pi = u[0, 0, 0], u[1, 0, 0], u[2, 0, 0], u[3, 0, 0]
    z = {3: -(pr + 1j * pi), 4: -(1j) * (pr + 1j * pi), 7: -(qr + 1j * qi), 8: -(1j) * (qr + 1j * qi)}[plane]
    got = r[2][5, 10] + 1j * r[3][5, 10]
    assert abs(got - z) < 1e-14


def test_d_term_and_dbdx_fold():
    """d q_x with complex d, and d_re's -(d bb/d rho) ee correction."""
    p = params(N=128, M=16)
    x = p.xmin + np.arange(p.N) * p.dx
    u = np.zeros((4, p.M, p.N))
    u[0] = 0.01 * x
    u[1] = -0.02 * x                               # q_x = 0.01 - 0.02 i
    coef = zero_coef(p)
    coef[6] = 1.0                                  # d_im = 1: i q_x
    r = O.rhs(p, coef, theta_of(p), u)
    qx = 0.01 - 0.02j
    assert abs((r[2][5, 40] + 1j * r[3][5, 40]) - 1j * qx) < 1e-13
    coef = zero_coef(p)
    coef[0] = 0.5 * x[None, :] / 100.0             # bb slope 0.005
    coef[2] = 2.0                                   # ee = 2
    u = np.zeros((4, p.M, p.N))
    u[0] = 0.01 * x                                 # q_x = 0.01, p = 0
    r = O.rhs(p, coef, theta_of(p), u)
    # rhs_p_re = ee*p_x + dre*q_x with dre = 0 - 0.005*2
    assert abs(r[2][5, 40] - (-0.005 * 2.0 * 0.01)) < 1e-13
    # rhs_q_re = -bb q_x + p
    assert abs(r[0][5, 40] - (-coef[0][5, 40] * 0.01)) < 1e-13


# ------------------------------------------------------------ boundaries ----
def test_theta_bc_even_chain_and_odd_zero():
    rng = np.random.default_rng(4)
    p = params(N=128, M=16, m=2)
    u = rng.standard_normal((4, p.M, p.N))
    v = u.copy()
    O.lib().orc_bc_theta(O.ctypes.byref(p), O._ptr(v))
    c = slice(3, p.N - 3)
    # each filled row makes the one-sided 2nd-order derivative vanish at it,
    # using the rows as they stand after the fill (so the order is pinned)
    for b in (2, 1, 0):
        assert np.allclose(-3 * v[:, b, c] + 4 * v[:, b + 1, c] - v[:, b + 2, c], 0, atol=1e-14)
    M = p.M
    for b in (M - 3, M - 2, M - 1):
        assert np.allclose(-3 * v[:, b, c] + 4 * v[:, b - 1, c] - v[:, b - 2, c], 0, atol=1e-14)
    assert np.array_equal(v[:, 3:M - 3], u[:, 3:M - 3])
    assert np.array_equal(v[:, :, :3], u[:, :, :3])            # rho edges untouched
    p1 = params(N=128, M=16, m=-3)
    w = u.copy()
    O.lib().orc_bc_theta(O.ctypes.byref(p1), O._ptr(w))
    for b in (0, 1, 2, M - 3, M - 2, M - 1):
        assert np.all(w[:, b, c] == 0.0)


def test_rho_bc_chain_after_theta():
    rng = np.random.default_rng(5)
    p = params(N=128, M=16, m=2)
    v = rng.standard_normal((4, p.M, p.N))
    O.bc(p, v)
    N = p.N
    for a in (2, 1, 0):
        assert np.allclose(-3 * v[:, :, a] + 4 * v[:, :, a + 1] - v[:, :, a + 2], 0, atol=1e-13)
    for a in (N - 3, N - 2, N - 1):
        assert np.allclose(-3 * v[:, :, a] + 4 * v[:, :, a - 1] - v[:, :, a - 2], 0, atol=1e-13)
    # corner rows were theta-filled first: rows 0..2 still satisfy the theta
    # relation on columns 3..4 that fed the rho fill
    assert np.allclose(-3 * v[:, 2, 3:5] + 4 * v[:, 3, 3:5] - v[:, 4, 3:5], 0, atol=1e-13)


# ------------------------------------------------------------- integrators --
@pytest.mark.parametrize("integrator", [0, 1])
def test_rk_amplification_polynomial(integrator):
    """p' = lambda p on constant fields: one step multiplies p by R(lambda dt)."""
    p = params(N=128, M=16, integrator=integrator)
    lam = (-0.3 + 0.7j) / p.dt * 0.5
    coef = zero_coef(p)
    coef[3] = -lam.real                           # rhs_p = -(a_re + i a_im) p = lam p
    coef[4] = -lam.imag
    u = np.zeros((4, p.M, p.N))
    u[2] = 1.0
    u[3] = 0.5
    p0 = 1.0 + 0.5j
    O.step(p, coef, theta_of(p), u)
    z = lam * p.dt
    R = 1 + z + z * z / 2 + z ** 3 / 6 + (z ** 4 / 24 if integrator == 1 else 0)
    got = u[2] + 1j * u[3]
    assert np.allclose(got, R * p0, rtol=1e-14, atol=1e-14)


def test_oracle_timing_build_matches_parity_build():
    from bh_ringdown_amd import teuk, workloads
    p = teuk.default_params(256, 32, -50.0, 50.0, 0.5, 2)
    _, r, th = teuk.grid(p)
    u0 = workloads.gaussian_pulse(p, 3, center=10.0)
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, nthreads=1)
    u1 = u0.copy()
    O.run(op, r, th, u1, 50)
    op.nthreads = 4
    u2 = u0.copy()
    psi = np.zeros((50, 2, p.M)); dedt = np.zeros(50)
    O.lib(timing=True).orc_run(O.ctypes.byref(op), O._ptr(r), O._ptr(th), O._ptr(u2), 50, None,
                               O._ptr(psi), None, O._ptr(dedt))
    assert np.linalg.norm(u2 - u1) / np.linalg.norm(u1) < 1e-12


@pytest.mark.parametrize("integrator", [0, 1])
def test_parity_build_thread_count_does_not_change_bits(integrator):
    """The parity build splits only point loops over threads: 1 and 5 threads
    give the same bits, with the source on."""
    from bh_ringdown_amd import teuk, workloads
    p = teuk.default_params(256, 32, -40.0, 60.0, 0.7, 2)
    steps = 30
    _, r, th = teuk.grid(p)
    u0 = workloads.gaussian_pulse(p, 5, center=10.0)
    rows = teuk.plunge_trajectory(p.a, 1e-2, 0.4, p.dt / 2, (2 * steps + 6) * p.dt / 2, seed=2,
                                  r0_over_isco=1.3)
    ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_])
           for x in teuk.trajectory_schedule(rows, steps, p.dt)]
    got = []
    for nt in (1, 5):
        op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=integrator, nthreads=nt)
        u = u0.copy()
        rc, psi, h, _ = O.run(op, r, th, u, steps, src_steps=ost)
        assert rc == 0
        got.append((u, psi, h))
    for a, b in zip(*got):
        assert np.array_equal(a, b)
