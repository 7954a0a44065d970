"""Host-side trajectory generator (teuk_traj_plunge: inspiral, Ori-Thorne
transition, plunge): its circular orbits, ISCO and geodesic rows checked
against closed forms and first integrals of Kerr geodesic motion; its
transition function against the reference's own table (GenOT/OT_X_vs_L.dat,
fixture ot_x_vs_l.npz); the whole worldline against the reference's own
plunge (QNM_Extraction/Examples/a0.7_thi000_thf090.0_n.traj, made by GenOT;
fixture examples_a07_traj.npz).  No GPU is used."""
import math
import os

import numpy as np
import pytest

from bh_ringdown_amd import teuk

HERE = os.path.dirname(os.path.abspath(__file__))


def kerr_isco_equatorial(a):
    """Bardeen-Press-Teukolsky (1972) prograde ISCO radius, M = 1."""
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = math.sqrt(3 * a * a + z1 * z1)
    return 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))


def R(r, a, E, L, Q):
    P = E * (r * r + a * a) - a * L
    D = r * r - 2 * r + a * a
    return P * P - D * (r * r + (L - a * E) ** 2 + Q)


def dt_dlambda(r, th, a, E, L):
    D = r * r - 2 * r + a * a
    P = E * (r * r + a * a) - a * L
    return -a * (a * E * math.sin(th) ** 2 - L) + (r * r + a * a) / D * P


@pytest.mark.parametrize("a,iota", [(0.0, 0.0), (0.0, 0.7), (0.5, 0.0), (0.9, 0.0), (0.9, math.radians(30))])
def test_start_is_circular_orbit_near_isco(a, iota):
    rows = teuk.plunge_trajectory(a, 1e-3, iota, 0.05, 1.0, seed=1, r0_over_isco=1.3)
    r0, th0, E, L, Q = rows[0, 1], rows[0, 2], rows[0, 7], rows[0, 8], rows[0, 9]
    if iota == 0.0 or a == 0.0:
        # equatorial (or Schwarzschild, any plane): closed-form ISCO
        assert r0 == pytest.approx(1.3 * kerr_isco_equatorial(a), abs=3e-6)
    if iota == 0.0:
        v = 1 / math.sqrt(r0)
        den = math.sqrt(1 - 3 * v * v + 2 * a * v ** 3)
        assert E == pytest.approx((1 - 2 * v * v + a * v ** 3) / den, abs=1e-10)
        assert L == pytest.approx(math.sqrt(r0) * (1 - 2 * a * v ** 3 + a * a * v ** 4) / den, abs=1e-9)
        assert th0 == pytest.approx(math.pi / 2, abs=1e-12)
    # R(r0) = R'(r0) = 0, and Q = Lz^2 tan^2(iota)
    h = 1e-6
    scale = E * E * r0 ** 4
    assert abs(R(r0, a, E, L, Q)) < 1e-10 * scale
    assert abs((R(r0 + h, a, E, L, Q) - R(r0 - h, a, E, L, Q)) / (2 * h)) < 1e-6 * scale
    assert Q == pytest.approx(L * L * math.tan(iota) ** 2, abs=1e-12)
    # the polar motion stays within the inclination: |cos theta| <= sin(iota)-ish
    assert abs(math.cos(th0)) <= math.sqrt(Q / (L * L + Q)) + 1e-12


def test_inclined_isco_lies_outside_equatorial():
    r_eq = teuk.plunge_trajectory(0.9, 1e-3, 0.0, 0.05, 1.0, r0_over_isco=1.3)[0, 1]
    r_30 = teuk.plunge_trajectory(0.9, 1e-3, math.radians(30), 0.05, 1.0, r0_over_isco=1.3)[0, 1]
    r_60 = teuk.plunge_trajectory(0.9, 1e-3, math.radians(60), 0.05, 1.0, r0_over_isco=1.3)[0, 1]
    assert r_eq < r_30 < r_60 < 1.3 * 6.0


def test_start_inside_the_transition_moves_out_to_it():
    """A start radius inside the transition region (here the ISCO itself)
    begins where the transition does, outside the ISCO; the worldline is
    finite for every spin and inclination the workloads draw."""
    r_isco = kerr_isco_equatorial(0.7)
    rows = teuk.plunge_trajectory(0.7, 1e-3, 0.0, 0.05, 1.0, r0_over_isco=1.0)
    assert r_isco + 0.02 < rows[0, 1] < r_isco + 0.5
    for a in (0.0, 0.3, 0.6, 0.631, 0.72, 0.9, 0.99):
        for iota in (0.0, math.radians(30), math.radians(80)):
            w = teuk.plunge_trajectory(a, 1e-3, iota, 0.5, 3000.0, seed=1, r0_over_isco=1.02)
            assert np.isfinite(w).all() and w[-1, 1] < 1 + math.sqrt(1 - a * a) + 1e-4, (a, iota)


def test_plunge_rows_follow_geodesic_first_integrals():
    a, dt_out = 0.9, 0.025
    rows = teuk.plunge_trajectory(a, 1e-3, math.radians(30), dt_out, 400.0, seed=7)
    t, r, th, rdot = rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3]
    E, L, Q = rows[:, 7], rows[:, 8], rows[:, 9]
    # uniform BL-time sampling starting at 0
    assert t[0] == 0.0
    assert np.allclose(np.diff(t), dt_out, rtol=0, atol=1e-9)
    # radiation reaction only removes energy and angular momentum
    assert np.all(np.diff(E) <= 1e-15) and np.all(np.diff(L) <= 1e-14)
    assert E[-1] < E[0]
    # it plunges: ends at the horizon, radius non-increasing once inside r0 - 0.1
    rh = 1 + math.sqrt(1 - a * a)
    assert r[-1] - rh < 1e-5
    k = np.argmax(r < r[0] - 0.1)
    assert k > 0 and np.all(np.diff(r[k:]) <= 1e-12)
    # once the constants are frozen (the plunge) the motion is a geodesic:
    # (dr/dl)^2 - R(r) = (dr/dt)^2 (dt/dl)^2 - R(r) is conserved (its value
    # is set by the transition's radial speed where the plunge starts)
    frozen = np.flatnonzero((E == E[-1]) & (r > rh + 0.05))
    assert frozen.size > 50
    inv = []
    for i in frozen[:: max(1, frozen.size // 40)]:
        tl = dt_dlambda(r[i], th[i], a, E[i], L[i])
        inv.append((rdot[i] * tl) ** 2 - R(r[i], a, E[i], L[i], Q[i]))
    inv = np.array(inv)
    scale = max(abs(R(r[frozen[0]], a, E[-1], L[-1], Q[-1])), np.max(np.abs(rdot[frozen] * 10.0)) ** 2, 1e-3)
    assert np.ptp(inv) < 1e-6 * scale


@pytest.mark.parametrize("a,eta,iota_deg", [(0.0926, 8.4e-3, 85.0), (0.9366, 7.4e-3, 84.0), (0.6, 4.1e-3, 85.0),
                                          (0.9, 1e-3, 30.0)])
def test_polar_motion_keeps_its_first_integral(a, eta, iota_deg):
    """Steep, fast inspirals (the serving stream draws inclinations up to 85
    degrees and mass ratios up to 1e-2) keep theta between its turning
    points, and on every row the polar speed satisfies Carter's first
    integral (sin(theta) dtheta/dl)^2 = Z(cos(theta)) with the row's
    constants, through the inspiral and transition (the constants change)
    as well as the plunge."""
    dt_out = 0.05
    rows = teuk.plunge_trajectory(a, eta, math.radians(iota_deg), dt_out, 600.0, seed=11)
    assert np.isfinite(rows).all()
    th, thdot, r = rows[:, 2], rows[:, 4], rows[:, 1]
    E, L, Q = rows[:, 7], rows[:, 8], rows[:, 9]
    margin = math.radians(90.0 - iota_deg) * 0.5
    assert th.min() > margin and th.max() < math.pi - margin
    rh = 1 + math.sqrt(1 - a * a)
    live = r > rh + 0.05
    z = np.cos(th[live])
    b = a * a * (1 - E[live] ** 2)
    Z = Q[live] - z * z * (b * (1 - z * z) + L[live] ** 2 + Q[live])
    tl = np.array([dt_dlambda(rr, tt, a, ee, ll) for rr, tt, ee, ll in zip(r[live], th[live], E[live], L[live])])
    lhs = (np.sin(th[live]) * thdot[live] * tl) ** 2
    assert np.max(np.abs(lhs - Z)) < 1e-9 * np.max(Q), np.max(np.abs(lhs - Z))


def test_transition_function_matches_genot_table():
    """X(L), the solution of X'' = -X^2 - L with X -> sqrt(-L), against every
    100th row of the reference's GenOT/OT_X_vs_L.dat over its whole range
    (L = -40 .. 3.4, X = 6.3 .. -2e6): measured within 4e-5 relative."""
    fx = np.load(os.path.join(HERE, "golden", "ot_x_vs_l.npz"))
    for L, want in zip(fx["L"], fx["X"]):
        x, dx = teuk.ot_transition(L)
        assert abs(x - want) <= 1e-4 * max(1.0, abs(want)), (L, x, want)
    # the differential equation at a few points (central differences)
    for L in (-20.0, -3.0, 0.0, 2.35):
        h = 1e-3
        x0, xp, xm = teuk.ot_transition(L)[0], teuk.ot_transition(L + h)[0], teuk.ot_transition(L - h)[0]
        assert abs((xp - 2 * x0 + xm) / h ** 2 + x0 * x0 + L) < 1e-4 * max(1.0, x0 * x0)
    with pytest.raises(RuntimeError):
        teuk.ot_transition(3.5)


def test_plunge_tracks_the_reference_example():
    """The a = 0.7 prograde equatorial worldline, mass ratio 1.3e-3 (the
    transition's time scale, eta B, is the one free parameter: GenOT's fluxes
    are adiabatic Teukolsky fluxes, these the quadrupole formula), aligned
    in time where r = 3: r(t) and dphi/dt stay within 0.04 M (measured 0.026)
    and 2 % (1.2 %) of the reference's GenOT plunge from r = 3.2 to 1.76,
    where it is at the horizon.  The transition matters: the model this
    generator had before (constants frozen at r_isco - 0.5, dr/dl relaxed
    onto the geodesic) ran 5 M late over the same stretch."""
    f = np.load(os.path.join(HERE, "golden", "examples_a07_traj.npz"))
    rows = teuk.plunge_trajectory(0.7, 1.3e-3, 0.0, 0.01, 8000.0, seed=0, r0_over_isco=1.3)
    t, r, phd = rows[:, 0], rows[:, 1], rows[:, 5]
    i3, j3 = np.argmin(np.abs(r - 3.0)), np.argmin(np.abs(f["r"] - 3.0))
    off = f["t"][j3] - t[i3]
    sel = (f["t"] > 2441.0) & (f["r"] > 1.76)
    assert sel.sum() > 2000
    rg = np.interp(f["t"][sel], t + off, r)
    pg = np.interp(f["t"][sel], t + off, phd)
    assert np.abs(rg - f["r"][sel]).max() < 0.04
    assert np.abs(pg / f["phidot"][sel] - 1).max() < 0.02


def test_seed_reproducible_and_distinct():
    a = teuk.plunge_trajectory(0.7, 1e-2, 0.5, 0.1, 50.0, seed=3)
    b = teuk.plunge_trajectory(0.7, 1e-2, 0.5, 0.1, 50.0, seed=3)
    c = teuk.plunge_trajectory(0.7, 1e-2, 0.5, 0.1, 50.0, seed=4)
    assert np.array_equal(a, b)
    assert a[0, 6] != c[0, 6]          # initial azimuth
    assert a[0, 2] != c[0, 2]          # initial polar phase
    assert a.shape == (501, 10)


def test_rejects_bad_arguments():
    with pytest.raises(RuntimeError):
        teuk.plunge_trajectory(1.0, 1e-3, 0.0, 0.1, 10.0)
    with pytest.raises(RuntimeError):
        teuk.plunge_trajectory(0.5, 0.0, 0.0, 0.1, 10.0)
    with pytest.raises(RuntimeError):
        teuk.plunge_trajectory(0.5, 1e-3, 0.0, 0.1, 10.0, r0_over_isco=0.9)


def test_traj_file_feeds_the_schedule(tmp_path):
    """Rows at dt/2 spacing drive the reference consumption pattern for the
    whole run (two rows per step)."""
    dt = 0.05
    rows = teuk.plunge_trajectory(0.9, 1e-3, math.radians(30), dt / 2, 20.0, seed=2)
    path = tmp_path / "plunge.traj"
    teuk.write_traj(path, rows)
    back = np.loadtxt(path)
    assert np.array_equal(back, rows)
    sched = teuk.trajectory_schedule(back, 390, dt)
    assert len(sched) == 390
    assert sched[-1].rp == back[2 * 389 + 3, 1] + 1e-6   # body.cu:112-116: row 3+2k
