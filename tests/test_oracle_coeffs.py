"""The oracle's coefficient fields (orc_coeffs, the restatement of
kern.cu:113-158, and the half-point set of the staggered scheme, the same
formulas at Old/kern.cu:93's half points) against Teukolsky's equation.

The nine fields define the first-order system of kern.cu:290-298,
    Psi_t = -bb Psi_x + Pi
    Pi_t  =  ee Pi_x + cc (Psi_yy + cot y Psi_y) + (d - bb_x ee) Psi_x - a Pi - v Psi
(a, d, v complex), which with Pi eliminated is the second-order operator
    Psi_tt = (ee - bb) Psi_tx + ee bb Psi_xx + cc (Psi_yy + cot y Psi_y)
             + (d - a bb) Psi_x - a Psi_t - v Psi .
Independently of the reference's expressions, sympy derives that operator
from the Teukolsky master equation for spin weight s = -2 in Boyer-Lindquist
coordinates (Teukolsky 1973, ApJ 185, 635, eq. 4.7) under
    psi = exp(i m phi~) r^3 Phi,      phi~ = phi + a Int dr / Delta,
    tau = t - h(rho),  h' = dr*/drho - 1,   r* = rho / Omega(rho),
the Kerr azimuth and the hyperboloidal layer of the compactified grid.  Each
of the six coefficient functions (tau-rho, rho-rho, theta, tau, rho, zeroth
order) is compared at grid points inside the layer and outside it, for
several spins and modes.  The operator fixes bb only up to the choice of
root; the choice is pinned by bb < 0 (the other root is positive)."""
import functools
import math

import mpmath as mp
import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O

S_WEIGHT = -2


@functools.lru_cache(maxsize=1)
def derived_operator():
    """Coefficients of d_tau^2 Phi = sum_k c_k (jet k of Phi), as mpmath
    functions of (r, rho, Omega, Omega', Omega'', theta, a, M, m)."""
    r, rho, om, om1, om2, th, a, M, m = sp.symbols("r rho Om Om1 Om2 th a M m", real=True)
    Q, Qt, Qx, Qy, Qtt, Qtx, Qxx, Qyy = sp.symbols("Q Qt Qx Qy Qtt Qtx Qxx Qyy")
    s, I = S_WEIGHT, sp.I
    delta = r**2 - 2 * M * r + a**2
    L = om - rho * om1                       # Omega^2 dr*/drho
    Lp = -rho * om2                          # dL/drho
    # J = drho/dr at fixed t, a function of (Omega, L, r); r depends on rho
    # through dr/drho = 1/J.
    Ls = sp.Symbol("Ls")
    Jf = om**2 * (r**2 + a**2) / (Ls * delta)
    J = Jf.subs(Ls, L)
    J_rho = (sp.diff(Jf, om) * om1 + sp.diff(Jf, Ls) * Lp + sp.diff(Jf, r) / Jf).subs(Ls, L)
    hp = L / om**2 - 1                       # dh/drho
    hpp = Lp / om**2 - 2 * L * om1 / om**3
    # Boyer-Lindquist derivatives of Phi in terms of its (tau, rho) jets
    P_r = J * (Qx - hp * Qt)
    P_rr = J * J_rho * (Qx - hp * Qt) + J**2 * (Qxx - 2 * hp * Qtx - hpp * Qt + hp**2 * Qtt)
    # psi / exp(i m phi) = F(r) Phi with F'/F = 3/r + i m a / Delta
    f1 = 3 / r + I * m * a / delta
    f2 = sp.diff(f1, r) + f1**2
    psi, psi_t, psi_tt = Q, Qt, Qtt
    psi_r = f1 * Q + P_r
    psi_rr = f2 * Q + 2 * f1 * P_r + P_rr
    psi_ph, psi_phph, psi_tph = I * m * Q, -(m**2) * Q, I * m * Qt
    sn, cs = sp.sin(th), sp.cos(th)
    T = (((r**2 + a**2) ** 2 / delta - a**2 * sn**2) * psi_tt
         + 4 * M * a * r / delta * psi_tph
         + (a**2 / delta - 1 / sn**2) * psi_phph
         - ((s + 1) * sp.diff(delta, r) * psi_r + delta * psi_rr)
         - (Qyy + cs / sn * Qy)
         - 2 * s * (a * (r - M) / delta + I * cs / sn**2) * psi_ph
         - 2 * s * (M * (r**2 - a**2) / delta - r - I * a * cs) * psi_t
         + (s**2 * cs**2 / sn**2 - s) * psi)
    T = sp.expand(T)
    jets = (Qtt, Qtx, Qxx, Qyy, Qy, Qt, Qx, Q)
    co = {str(k): T.coeff(k) for k in jets}
    rest = sp.expand(T - sum(T.coeff(k) * k for k in jets))
    assert rest == 0
    for v in co.values():
        assert not v.has(*jets)
    args = (r, rho, om, om1, om2, th, a, M, m)
    return {k: sp.lambdify(args, v, "mpmath") for k, v in co.items()}


def omega(p, x):
    """The compactification of kern.cu:116-126 (the grid's definition, not an
    operator coefficient): Omega and its first two derivatives."""
    rin = p.xmin + (p.xmax - p.xmin) * 0.64
    outer = p.xmax + 1e-6
    if x < rin:
        return 1.0, 0.0, 0.0
    w = outer - rin
    return 1.0 - ((x - rin) / w) ** 4, -4.0 * (x - rin) ** 3 / w**4, -12.0 * (x - rin) ** 2 / w**4


def want_at(p, r, x, y):
    f = derived_operator()
    mp.mp.dps = 50
    om, om1, om2 = omega(p, x)
    args = [mp.mpf(float(v)) for v in (r, x, om, om1, om2, y, p.a, p.mass, p.m)]
    c = {k: mp.mpc(fn(*args)) for k, fn in f.items()}
    n = -c["Qtt"]
    return {"Atx": c["Qtx"] / n, "Axx": c["Qxx"] / n, "Ayy": c["Qyy"] / n, "Ay": c["Qy"] / n,
            "Bt": c["Qt"] / n, "Bx": c["Qx"] / n, "C": c["Q"] / n}


def got_at(coef, k, j, y):
    """The operator's coefficients from the fields, each with the magnitude
    of the terms it is formed from (the scale of its rounding error)."""
    bb, cc, ee = coef[0, k, j], coef[1, k, j], coef[2, k, j]
    a = complex(coef[3, k, j], coef[4, k, j])
    d = complex(coef[5, k, j], coef[6, k, j])
    v = complex(coef[7, k, j], coef[8, k, j])
    cot = np.cos(y) / np.sin(y)
    return {"Atx": (ee - bb, abs(ee) + abs(bb)), "Axx": (ee * bb, abs(ee * bb)),
            "Ayy": (cc, abs(cc)), "Ay": (cc * cot, abs(cc * cot)),
            "Bt": (-a, abs(a)), "Bx": (d - a * bb, abs(d) + abs(a * bb)), "C": (-v, abs(v))}


def columns(p, r, x):
    """Sample columns inside the layer and outside it, away from the horizon,
    where Delta loses all digits in double precision."""
    delta = r * r - 2.0 * p.mass * r + p.a * p.a
    ok = np.nonzero(delta > 1e-4 * r * r)[0]
    rin = p.xmin + (p.xmax - p.xmin) * 0.64
    inner = ok[x[ok] < rin]
    layer = ok[x[ok] >= rin]
    pick = lambda v: v[np.linspace(0, len(v) - 1, 5).round().astype(int)]
    return np.concatenate([pick(inner), pick(layer), [p.N - 2]])


def check(p, coef, x, r, theta, rows, rtol=1e-11):
    worst = 0.0
    cols = columns(p, r, x)
    for k in rows:
        for j in cols:
            want = want_at(p, r[j], x[j], theta[k])
            got = got_at(coef, k, j, theta[k])
            for key, w in want.items():
                w = complex(w)
                g, scale = got[key]
                err = abs(g - w) / max(abs(w), scale, 1e-300)
                worst = max(worst, err)
                assert err < rtol, (p.a, p.m, k, j, key, g, w)
    return worst


@pytest.mark.parametrize("a,m", [(0.0, 2), (0.3, 1), (0.6, -3), (0.9, 2), (0.99, 4), (0.5, -1)])
def test_high_order_coefficients_solve_teukolsky(a, m):
    p = O.make_params(256, 16, -100.0, 300.0, a, m)
    x, r_c, theta = O.grid(p)
    coef = O.coeffs(p, theta, r_c)
    check(p, coef, x, r_c, theta, rows=(3, 5, 8, 12))
    assert np.all(coef[:, :3] == 0.0) and np.all(coef[:, p.M - 3:] == 0.0)
    # the root of kern.cu:147
    assert np.all(coef[0, 3:p.M - 3] < 0.0)


@pytest.mark.parametrize("a,m", [(0.6, 2), (0.9, -3)])
def test_half_point_coefficients_solve_teukolsky(a, m):
    """The staggered scheme's half-point set at rho = Xmin + (j - 1/2) dx with
    the half-point radii, and its grid-point set."""
    p = O.make_params(256, 16, -100.0, 300.0, a, m, integrator=2)
    _, r_c, theta = O.grid(p)
    x_h, r_h = O.grid_half(p.N, p.xmin, p.xmax, p.mass, p.a)
    ch, cc = O.coeffs_so2(p, theta, r_h, r_c)
    check(p, ch, x_h, r_h, theta, rows=(3, 6, 12))
    x_c = p.xmin + np.arange(p.N) * p.dx
    check(p, cc, x_c, r_c, theta, rows=(3, 6, 12))


@pytest.mark.parametrize("plant", ["conjugate a", "d without bb B^tau"])
def test_planted_coefficient_error_is_caught(plant):
    """A sign slip (Im a flipped) or a dropped term (d = B^rho, missing
    -bb B^tau) on one row of the oracle's output is found."""
    p = O.make_params(256, 16, -100.0, 300.0, 0.7, 2)
    x, r_c, theta = O.grid(p)
    bad = O.coeffs(p, theta, r_c)
    k = 5
    if plant == "conjugate a":
        bad[4, k] = -bad[4, k]
    else:
        bad[5, k] = bad[5, k] - bad[0, k] * bad[3, k]
        bad[6, k] = bad[6, k] - bad[0, k] * bad[4, k]
    with pytest.raises(AssertionError):
        check(p, bad, x, r_c, theta, rows=(k,))


@pytest.mark.parametrize("a,m", [(0.0, 2), (0.3, 1), (0.6, -3), (0.9, 2), (0.99, 4), (0.5, -1)])
def test_coefficients_near_the_horizon(a, m):
    """The innermost columns, where Delta = r^2 - 2 M r + a^2 keeps fewer
    digits than r (Delta < 1e-4 r^2): the same comparison with Teukolsky's
    operator evaluated exactly (mpmath) at the same double-precision r.  The
    double evaluation loses about eps r^2 / Delta to the cancellation in
    Delta (measured up to 4e-4 where r^2 / Delta = 1e13), so the tolerance is
    1e-15 r^2 / Delta + 1e-11 -- a wrong formula misses by O(1) at every
    column.  Columns where the grid's Newton inversion was clamped to r = r+
    (Delta = 0 in double; the reference's own inversion gives NaN there) are
    left out."""
    p = O.make_params(256, 16, -100.0, 300.0, a, m)
    x, r, theta = O.grid(p)
    coef = O.coeffs(p, theta, r)
    rplus = 1.0 + math.sqrt((1.0 - a) * (1.0 + a))
    mp.mp.dps = 40
    cols = []
    for j in range(p.N):
        dl = mp.mpf(r[j]) ** 2 - 2 * mp.mpf(r[j]) + mp.mpf(a) ** 2
        if dl > 0 and dl < 1e-4 * r[j] ** 2:
            cols.append((j, dl))
    assert len(cols) >= 20
    checked = 0
    for j, dl in cols[::max(1, len(cols) // 10)]:
        bound = 1e-15 * r[j] ** 2 / float(dl) + 1e-11
        for k in (3, 12):
            want = want_at(p, r[j], x[j], theta[k])
            got = got_at(coef, k, j, theta[k])
            for key, w in want.items():
                w = complex(w)
                g, scale = got[key]
                err = abs(g - w) / max(abs(w), scale, 1e-300)
                assert err < bound, (a, m, j, k, key, err, bound)
        checked += 1
    assert checked >= 10
    assert r.min() >= rplus
