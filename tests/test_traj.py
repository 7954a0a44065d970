"""Host-side plunge trajectory generator (teuk_traj_plunge): its circular
orbits, ISCO and geodesic rows checked against closed forms and first
integrals of Kerr geodesic motion.  No GPU is used."""
import math

import numpy as np
import pytest

from bh_ringdown_amd import teuk


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
    rows = teuk.plunge_trajectory(a, 1e-3, iota, 0.05, 1.0, seed=1, r0_over_isco=1.0)
    r0, th0, E, L, Q = rows[0, 1], rows[0, 2], rows[0, 7], rows[0, 8], rows[0, 9]
    if iota == 0.0 or a == 0.0:
        # equatorial (or Schwarzschild, any plane): closed-form ISCO
        assert r0 == pytest.approx(kerr_isco_equatorial(a), abs=2e-6)
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
    r_eq = teuk.plunge_trajectory(0.9, 1e-3, 0.0, 0.05, 1.0, r0_over_isco=1.0)[0, 1]
    r_30 = teuk.plunge_trajectory(0.9, 1e-3, math.radians(30), 0.05, 1.0, r0_over_isco=1.0)[0, 1]
    r_60 = teuk.plunge_trajectory(0.9, 1e-3, math.radians(60), 0.05, 1.0, r0_over_isco=1.0)[0, 1]
    assert r_eq < r_30 < r_60 < 6.0


def test_plunge_rows_follow_geodesic_first_integrals():
    a, dt_out = 0.9, 0.025
    rows = teuk.plunge_trajectory(a, 1e-3, math.radians(30), dt_out, 400.0, seed=7)
    t, r, th, rdot = rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3]
    E, L, Q = rows[:, 7], rows[:, 8], rows[:, 9]
    # uniform BL-time sampling starting at 0
    assert t[0] == 0.0
    assert np.allclose(np.diff(t), dt_out, rtol=0, atol=1e-9)
    # radiation reaction only removes energy and angular momentum
    assert np.all(np.diff(E) <= 0) and np.all(np.diff(L) <= 0)
    assert E[-1] < E[0]
    # it plunges: ends at the horizon, radius non-increasing once inside r0 - 0.1
    rh = 1 + math.sqrt(1 - a * a)
    assert r[-1] - rh < 1e-5
    k = np.argmax(r < r[0] - 0.1)
    assert k > 0 and np.all(np.diff(r[k:]) <= 1e-12)
    # once the constants are frozen and the velocity has relaxed onto the
    # geodesic (a few M): (dr/dt)^2 (dt/dlambda)^2 = R(r)
    t_freeze = t[np.argmax(E == E[-1])]
    frozen = np.flatnonzero((E == E[-1]) & (t > t_freeze + 20.0) & (r > rh + 0.05))
    assert frozen.size > 50
    for i in frozen[:: max(1, frozen.size // 40)]:
        tl = dt_dlambda(r[i], th[i], a, E[i], L[i])
        lhs = (rdot[i] * tl) ** 2
        rhs = R(r[i], a, E[i], L[i], Q[i])
        assert lhs == pytest.approx(rhs, rel=1e-6, abs=1e-9)


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
