"""The oracle's staggered second-order scheme (TD_GPU_code/Old) against what
the mathematics fixes: the two-step update's amplification polynomial, the
exactness of its differences on low-degree polynomials (which also fixes the
direction of the staggering), the edge extrapolation and the order of the
edge fills, the half-point radii, and the coefficient sets."""
import numpy as np
import pytest

from oracle import oracle as O


def params(N=64, M=16, m=2, dt=0.01, xmin=-20.0, xmax=40.0, a=0.5):
    return O.make_params(N, M, xmin, xmax, a, m, dt=dt, integrator=2, nthreads=1)


def theta_of(p):
    return np.arange(p.M) * (O.PI15 / p.M)


def rho_of(p):
    return p.xmin + np.arange(p.N) * p.dx


def coef_planes(p, **planes):
    """Constant coefficient planes (bb, cc, ee, a_re, a_im, d_re, d_im, v_re, v_im)."""
    names = ["bb", "cc", "ee", "are", "aim", "dre", "dim", "vre", "vim"]
    c = np.zeros((9, p.M, p.N))
    for k, v in planes.items():
        c[names.index(k)] = v
    return c


@pytest.mark.parametrize("lam", [-0.7 + 2.0j, 0.3 - 1.1j])
def test_two_step_amplification_polynomial(lam):
    """Constant data, p' = lam p, q' = p: the half step then the full step
    give p(dt) = (1 + z + z^2/2) p and q(dt) = q + dt p + dt^2/2 lam p."""
    p = params(dt=0.05)
    c = coef_planes(p, are=-lam.real, aim=-lam.imag)
    u = np.zeros((4, p.M, p.N))
    u[0], u[2] = 1.0, 1.0
    O.step_so2(p, c, c, theta_of(p), u)
    z = lam * p.dt
    pz = u[2] + 1j * u[3]
    qz = u[0] + 1j * u[1]
    assert np.allclose(pz, 1 + z + z * z / 2, rtol=0, atol=1e-14)
    assert np.allclose(qz, 1 + p.dt + p.dt ** 2 / 2 * lam, rtol=0, atol=1e-14)


def test_rho_differences_are_staggered_the_reference_way():
    """p' = p_rho with p = rho^2: the half step differences across the cell
    to the left of each point ((p[a] - p[a-1])/dx = 2 rho_{a-1/2}), the full
    step across the cell to the right of the half point, so
    p(dt) = rho^2 + dt (2 rho + dt) exactly in the interior; a difference
    taken the other way round misses by 2 dx dt."""
    p = params(dt=0.02)
    c = coef_planes(p, ee=1.0)
    x = rho_of(p)
    u = np.zeros((4, p.M, p.N))
    u[2] = x ** 2
    O.step_so2(p, c, c, theta_of(p), u)
    want = x ** 2 + p.dt * (2 * x + p.dt)
    inner = u[2][1:-1, 1:-1]
    assert np.allclose(inner, np.broadcast_to(want[1:-1], inner.shape), rtol=1e-13, atol=1e-12)


def test_theta_operator_exact_on_quadratics():
    """p' = q_thth + cot q_th with q = 1 + 0.3 th - 0.2 th^2 (constant in rho):
    centred second-order differences are exact on quadratics, so after one
    step p = dt (2 * (-0.2) + cot th (0.3 - 0.4 th)) on rows 2..M-3 (rows 1
    and M-2 see the edge rows, which the half step re-extrapolated)."""
    p = params(dt=0.01, M=32)
    c = coef_planes(p, cc=1.0)
    th = theta_of(p)
    u = np.zeros((4, p.M, p.N))
    u[0] = (1 + 0.3 * th - 0.2 * th ** 2)[:, None]
    O.step_so2(p, c, c, th, u)
    b = np.arange(2, p.M - 2)
    want = p.dt * (-0.4 + np.cos(th[b]) / np.sin(th[b]) * (0.3 - 0.4 * th[b]))
    got = u[2][2:-2, 1:-1]
    assert np.allclose(got, np.broadcast_to(want[:, None], got.shape), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("edge", ["theta_lo", "theta_hi", "rho_lo", "rho_hi"])
def test_edges_are_even_quadratic_extrapolation(edge):
    """With all coefficients zero and Pi = 0 nothing moves inside; each edge
    line becomes the value at that line of the quadratic that is even about
    it through the next two lines: data c + k (distance to the edge)^2 give
    back c there."""
    p = params(m=2)
    c = coef_planes(p)
    b = np.arange(p.M)[:, None] * np.ones((1, p.N))
    a = np.ones((p.M, 1)) * np.arange(p.N)[None, :]
    dist = {"theta_lo": b, "theta_hi": (p.M - 1) - b, "rho_lo": a, "rho_hi": (p.N - 1) - a}[edge]
    u = np.zeros((4, p.M, p.N))
    u[0] = 2.5 + 0.01 * dist ** 2
    O.step_so2(p, c, c, theta_of(p), u)
    line = {"theta_lo": u[0][0, 1:-1], "theta_hi": u[0][-1, 1:-1],
            "rho_lo": u[0][1:-1, 0], "rho_hi": u[0][1:-1, -1]}[edge]
    assert np.allclose(line, 2.5, rtol=0, atol=1e-13)


def test_theta_rows_before_rho_columns_and_odd_m_zero():
    """The corner takes the rho extrapolation of the theta-filled edge row
    (Old/body.cu:191-197 runs boundary2a before boundary2b): separable data
    g(b) h(a), each even quadratic about the corner, return g(0) h(0) there
    even when the old edge rows hold garbage.  Odd m zeroes the theta rows."""
    p = params(m=2)
    c = coef_planes(p)
    b = np.arange(p.M, dtype=float)
    a = np.arange(p.N, dtype=float)
    g = 3.0 + 0.02 * b ** 2
    h = 0.5 + 0.001 * a ** 2
    u = np.zeros((4, p.M, p.N))
    u[1] = np.outer(g, h)
    u[1][0, :] = 1e3
    u[1][:, 0] = -1e3
    O.step_so2(p, c, c, theta_of(p), u)
    assert abs(u[1][0, 0] - 3.0 * 0.5) < 1e-12
    q = params(m=3)
    v = np.ones((4, q.M, q.N))
    O.step_so2(q, c, c, theta_of(q), v)
    assert np.all(v[:, 0, 1:-1] == 0.0) and np.all(v[:, -1, 1:-1] == 0.0)


@pytest.mark.parametrize("a", [0.0, 0.5, 0.95])
def test_half_point_radii_solve_the_tortoise_relation(a):
    """r_h satisfies r*(r_h) = rho_h / Omega(rho_h) at rho_h = Xmin + (l - 1/2) dx."""
    N, xmin, xmax, mass = 512, -100.0, 300.0, 1.0
    x, r = O.grid_half(N, xmin, xmax, mass, a)
    dx = (xmax - xmin) / N
    assert np.allclose(x, xmin + (np.arange(N) - 0.5) * dx, rtol=0, atol=1e-12)
    rin, outer = xmin + 0.64 * (xmax - xmin), xmax + 1e-6
    om = np.where(x < rin, 1.0, 1.0 - ((x - rin) / (outer - rin)) ** 4)
    rp = mass + np.sqrt(mass * mass - a * a)
    rm = mass - np.sqrt(mass * mass - a * a)
    with np.errstate(divide="ignore", invalid="ignore"):   # points clamped to r+ give log(0)
        if a == 0.0:
            rstar = r + 2 * mass * np.log(r / (2 * mass) - 1)
        else:
            rstar = (r + (rp * rp + a * a) / (rp - rm) * np.log(r / rp - 1)
                     - (rm * rm + a * a) / (rp - rm) * np.log(r / rm - 1))
    ok = r > rp + 1e-4     # nearer the horizon r* amplifies the last bit of r past 1e-9
    assert ok.sum() > N // 2
    assert np.allclose(rstar[ok], (x / om)[ok], rtol=1e-10, atol=1e-10)


def test_coefficient_sets():
    """Centre set = the HighOrder fields (same formulas, kern.cu:132-158 and
    Old/kern.cu:143-181) on rows 3..M-4, and set on rows 1..M-2 only; the
    half set at column j lies between centre columns j-1 and j to O(dx^2) (a
    half-point shift the other way misses by O(dx))."""
    p = params(N=256, M=16, a=0.7, xmin=-50.0, xmax=50.0)
    th = theta_of(p)
    _, r_c, _ = O.grid(O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m))
    _, r_h = O.grid_half(p.N, p.xmin, p.xmax, p.mass, p.a)
    ch, cc = O.coeffs_so2(p, th, r_h, r_c)
    hi = O.coeffs(O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m), th, r_c)
    assert np.array_equal(cc[:, 3:-3], hi[:, 3:-3])
    assert np.all(cc[:, 0] == 0) and np.all(cc[:, -1] == 0) and np.all(ch[:, 0] == 0)
    assert np.all(cc[0, 1:-1] != 0) and np.all(ch[0, 1:-1] != 0)
    # inner region (Omega = 1): smooth in rho; compare to the centre average
    cols = np.arange(8, 120)
    mid = 0.5 * (cc[:, 1:-1, cols - 1] + cc[:, 1:-1, cols])
    scale = np.abs(cc[:, 1:-1, cols]).max(axis=(1, 2), keepdims=True) + 1e-300
    err = np.abs(ch[:, 1:-1, cols] - mid) / scale
    shifted = np.abs(ch[:, 1:-1, cols] - 0.5 * (cc[:, 1:-1, cols] + cc[:, 1:-1, cols + 1])) / scale
    assert err.max() < 1e-3
    assert shifted.max() > 10 * err.max()
