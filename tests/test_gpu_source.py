"""GPU parity with the particle source switched on: libteuk's set-up kernel
(T_4 coefficients per step), patch kernel and stage kernels against the
oracle, through the C ABI.  Bit-for-bit: both sides evaluate the same
expressions in the same order without contraction, with the same host libm
values for the few transcendental inputs."""
import math

import numpy as np
import pytest

from bh_ringdown_amd import teuk, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def inspiral_rows(n, dt, r0=8.0, rdot=-0.004, th0=1.5, thamp=0.05, om=0.04):
    """A smooth synthetic worldline sampled at dt/2 (the .traj spacing the
    reference consumes, body.cu:112-113): columns t r th rdot thdot phidot
    phi E Lz Q."""
    t = np.arange(n) * (dt / 2)
    rows = np.zeros((n, 10))
    rows[:, 0] = t
    rows[:, 1] = r0 + rdot * t
    rows[:, 2] = th0 + thamp * np.sin(0.02 * t)
    rows[:, 3] = rdot
    rows[:, 4] = thamp * 0.02 * np.cos(0.02 * t)
    rows[:, 5] = om * (1 + 0.001 * t)
    rows[:, 6] = om * (t + 0.0005 * t * t)
    rows[:, 7], rows[:, 8], rows[:, 9] = 0.95, 3.2, 0.0
    return rows


def assert_same_bits(got, want):
    """Bit-for-bit equality; on failure name the first differing cells."""
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    if same.all():
        return
    bad = np.argwhere(~same)
    lines = [f"{bad.shape[0]} of {got.size} values differ; first:"]
    for idx in bad[:8]:
        t = tuple(idx)
        lines.append(f"  {t}: got {got[t]!r} want {want[t]!r}")
    raise AssertionError("\n".join(lines))


def run_both(p, steps, rows):
    _, r_c, th = teuk.grid(p)
    sched = teuk.trajectory_schedule(rows, steps, p.dt)
    with teuk.Solver(p, r_c, th) as s:
        s.set_trajectory(rows, steps)
        s.step(steps)
        s.sync()
        ug, out, psig = s.get_fields(), s.output(), s.psi()
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=p.integrator,
                       j_extract=p.j_extract)
    ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
    uo = np.zeros((4, p.M, p.N))
    rc, psio, ho, dedto = O.run(op, r_c, th, uo, steps, src_steps=ost)
    assert rc == 0
    return ug, uo, out, psig, psio, ho


@pytest.mark.parametrize("a,m,integrator", [(0.5, 2, 0), (0.9, -3, 0), (0.7, 1, 1)])
def test_source_run_bitwise(a, m, integrator):
    steps = 300
    p = teuk.default_params(1024, 32, -50.0, 100.0, a, m, dt_over_dx=0.5)
    p.T = steps * p.dt * (1 + 1e-12)
    p.integrator = integrator
    p.j_extract = 900
    rows = inspiral_rows(2 * steps + 4, p.dt)
    print(f"source case a={a} m={m} integ={integrator} steps={steps}", flush=True)
    ug, uo, out, psig, psio, ho = run_both(p, steps, rows)
    assert np.any(uo != 0.0)               # the source did something
    assert_same_bits(ug, uo)
    assert np.array_equal(psig, psio[out["k"]])
    assert np.array_equal(out["h"], ho[out["k"]])


def oracle_coeffs(p, sched):
    """T_4 coefficients per step from the oracle, from the same schedule."""
    out = []
    for st in sched:
        state = {k: getattr(st, k) for k in ("E", "lz", "rp", "drdt", "d2rdt2", "dthdt", "d2thdt2", "dphidt",
                                             "d2phidt2", "tp", "phip")}
        state.update(mass=p.mass, a=p.a, nmu=1.0, mm=p.m)
        out.append(O.t4_coeffs(state))
    return np.array(out)


@pytest.mark.parametrize("case", ["interior", "edge"])
def test_source_coefficients_bitwise(case):
    """The device's per-step T_4 coefficients (one thread per step) against
    the oracle's, bit for bit, for every step of the schedule."""
    steps = 60
    if case == "interior":
        p = teuk.default_params(1024, 32, -50.0, 100.0, 0.9, -3, dt_over_dx=0.5)
        rows = inspiral_rows(2 * steps + 4, p.dt)
    else:
        p = teuk.default_params(512, 32, -10.0, 90.0, 0.6, 2, dt_over_dx=0.5)
        rows = inspiral_rows(2 * steps + 4, p.dt, r0=1.85, rdot=0.0, th0=0.2, thamp=0.0)
    p.T = steps * p.dt * (1 + 1e-12)
    sched = teuk.trajectory_schedule(rows, steps, p.dt)
    with teuk.Solver(p) as s:
        s.set_trajectory(rows, steps)
        got = s.source_coeffs(0, steps)
    assert_same_bits(got, oracle_coeffs(p, sched))


@pytest.mark.parametrize("case", ["interior", "edge"])
def test_source_patch_bitwise(case):
    """The device's source patch (Gaussian deltas, post-factors, clipping)
    against the oracle's source array at several steps, bit for bit."""
    steps = 60
    if case == "interior":
        p = teuk.default_params(1024, 32, -50.0, 100.0, 0.9, -3, dt_over_dx=0.5)
        rows = inspiral_rows(2 * steps + 4, p.dt)
    else:
        p = teuk.default_params(512, 32, -10.0, 90.0, 0.6, 2, dt_over_dx=0.5)
        rows = inspiral_rows(2 * steps + 4, p.dt, r0=1.85, rdot=0.0, th0=0.2, thamp=0.0)
    p.T = steps * p.dt * (1 + 1e-12)
    _, r_c, th = teuk.grid(p)
    sched = teuk.trajectory_schedule(rows, steps, p.dt)
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt)
    with teuk.Solver(p, r_c, th) as s:
        s.set_trajectory(rows, steps)
        for k in (1, 2, 17, 40, 59):
            patch, (r0, c0) = s.source_patch(k)
            st = O.SrcStep(*[getattr(sched[k], f) for f, _ in teuk.SourceStep._fields_])
            full = O.source(op, th, r_c, st)
            want = np.zeros_like(patch)
            rr = np.arange(p.ptsy) + r0
            cc = np.arange(p.ptsx) + c0
            ok_r = (rr >= 0) & (rr < p.M)
            ok_c = (cc >= 0) & (cc < p.N)
            want[:, np.ix_(ok_r, ok_c)[0], np.ix_(ok_r, ok_c)[1]] = full[:, rr[ok_r]][:, :, cc[ok_c]]
            print(f"patch case={case} step={k}", flush=True)
            assert_same_bits(patch, want)


@pytest.mark.parametrize("steps", [2, 60])   # step 0 has turn-on factor 0
def test_patch_at_domain_edges(steps):
    """Particle close to the inner rho edge and the theta axis: the patch is
    clipped by the reference's guards (index > 0, row < M) on both sides."""
    p = teuk.default_params(512, 32, -10.0, 90.0, 0.6, 2, dt_over_dx=0.5)
    p.T = steps * p.dt * (1 + 1e-12)
    # r = 1.85 (r+ = 1.8): tortoise radius about -6.3, patch origin below column 0
    rows = inspiral_rows(2 * steps + 4, p.dt, r0=1.85, rdot=0.0, th0=0.2, thamp=0.0)
    print(f"edge case steps={steps}", flush=True)
    ug, uo, *_ = run_both(p, steps, rows)
    assert_same_bits(ug, uo)


@pytest.mark.parametrize("N,M", workloads.HIGH_SPIN_GRIDS)
def test_high_spin_ladder_full_size(N, M):
    """The a=0.99 resolution ladder on rho in [-600,1600] at its full grid
    sizes (up to 65536 x 128), classic RK4, a Gaussian pulse plus a seeded
    plunge source: 8 steps against the oracle, bit for bit."""
    steps = 8
    p = workloads.high_spin(N, M)
    p.T = steps * p.dt * (1 + 1e-12)
    rows = workloads.plunge_rows(p, steps, seed=4)
    _, r_c, th = teuk.grid(p)
    u0 = workloads.gaussian_pulse(p, seed=6, center=200.0, width=10.0)
    print(f"high-spin case N={N} M={M} steps={steps}", flush=True)
    with teuk.Solver(p, r_c, th) as s:
        s.set_fields(u0)
        s.set_trajectory(rows, steps)
        s.step(steps)
        s.sync()
        ug, out, psig = s.get_fields(), s.output(), s.psi()
    sched = teuk.trajectory_schedule(rows, steps, p.dt)
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=p.integrator,
                       j_extract=p.j_extract)
    ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
    uo = u0.copy()
    rc, psio, ho, _ = O.run(op, r_c, th, uo, steps, src_steps=ost)
    assert rc == 0
    assert_same_bits(ug, uo)
    assert np.array_equal(psig, psio[out["k"]])
    assert np.array_equal(out["h"], ho[out["k"]])


@pytest.mark.parametrize("N,M", workloads.HIGH_SPIN_GRIDS[-1:])
def test_high_spin_affine_at_full_size(N, M):
    """With the source on, one step is an affine map u -> A u + s, so
    F(u1 + 2 u2) + 2 F(0) = F(u1) + 2 F(u2) after any number of steps; checked
    at 65536 x 128 over 40 RK4 steps."""
    steps = 40
    p = workloads.high_spin(N, M)
    p.T = steps * p.dt * (1 + 1e-12)
    rows = workloads.plunge_rows(p, steps, seed=4)
    _, r_c, th = teuk.grid(p)
    u1 = workloads.gaussian_pulse(p, seed=21, center=100.0, width=10.0)
    u2 = workloads.gaussian_pulse(p, seed=22, center=600.0, width=10.0)

    def F(u):
        with teuk.Solver(p, r_c, th) as s:
            s.set_fields(u)
            s.set_trajectory(rows, steps)
            s.step(steps)
            s.sync()
            return s.get_fields()

    print(f"high-spin affine case N={N} M={M} steps={steps}", flush=True)
    lhs = F(u1 + 2.0 * u2) + 2.0 * F(np.zeros_like(u1))
    rhs = F(u1) + 2.0 * F(u2)
    assert np.linalg.norm(lhs - rhs) <= 1e-13 * np.linalg.norm(rhs)
