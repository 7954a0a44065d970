"""Failure detection and source switching through the C ABI.

* The whole-field sentinel (teuk_status): a non-finite or runaway value
  anywhere in one solve's fields fails that solve after the chunk of steps it
  appears in, far from the extraction column too; in a batch only that member
  fails and its neighbours end bit for bit where they end alone.  A NaN or a
  huge number planted in the fields is ordinary data to the kernels (no field
  value is ever used as an index).
* Switching the source off (teuk_set_source with no steps) takes effect for
  every later step, however the steps are split into calls (captured step
  graphs are rebuilt), and the fields equal the oracle's with the source on
  for the first steps and off after.
"""
import numpy as np
import pytest

from bh_ringdown_amd import teuk, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def params(N, M, steps, a=0.6, m=2, integrator=1):
    p = teuk.default_params(N, M, -40.0, 60.0, a, m, dt_over_dx=0.5)
    p.T = steps * p.dt * (1 + 1e-12)
    p.integrator = integrator
    p.Fr = 4
    return p


def solve_alone(p, u0, rows, steps):
    with teuk.Solver(p) as s:
        s.set_fields(u0)
        if rows is not None:
            s.set_trajectory(rows, steps)
        s.step(steps)
        assert s.status() == teuk.TEUK_OK
        s.sync()
        return s.get_fields(), s.output()["h"], s.psi()


@pytest.mark.parametrize("value", [np.nan, np.inf, 1e70])
@pytest.mark.parametrize("integrator", [1, 0, 2])
def test_sentinel_catches_a_bad_value_far_from_extraction(value, integrator):
    """One bad value near the inner boundary: the solve fails at its first
    chunk, long before anything could reach the extraction column."""
    p = params(256, 32, 64, integrator=integrator)
    u0 = workloads.gaussian_pulse(p, seed=2, center=10.0)
    u0[0, 16, 20] = value                 # q_re, theta row 16, rho column 20
    print(f"sentinel integrator={integrator} value={value}", flush=True)
    with teuk.Solver(p) as s:
        s.set_fields(u0)
        s.step(8)
        assert s.status() == teuk.TEUK_EDIVERGED
        with pytest.raises(teuk.TeukError) as e:
            s.sync()
        assert e.value.code == teuk.TEUK_EDIVERGED
        psi = s.psi()
        assert np.all(np.isfinite(psi))   # the extraction column has not seen it


def test_sentinel_quiet_on_healthy_runs():
    p = params(256, 32, 96)
    u0 = workloads.gaussian_pulse(p, seed=3, center=10.0)
    rows = teuk.plunge_trajectory(p.a, 1e-2, 0.3, p.dt / 2, (2 * 96 + 6) * p.dt / 2, seed=3, r0_over_isco=1.3)
    with teuk.Solver(p) as s:
        s.set_fields(u0)
        s.set_trajectory(rows, 96)
        s.step(96)
        s.sync()
        assert s.status() == teuk.TEUK_OK


@pytest.mark.parametrize("integrator", [1, 0])
def test_failed_member_retired_neighbours_bitwise_alone(integrator):
    """Three members, the middle one carries a NaN: after the first chunk the
    batch reports a failure, only that member's status says so, and once it
    is removed the others run to their end bit for bit as alone."""
    steps = [70, 50, 90]
    specs = []
    for i, st in enumerate(steps):
        p = params(256 + 128 * i, 32 if i != 1 else 16, st, a=0.3 + 0.2 * i, m=2 - i, integrator=integrator)
        u0 = workloads.gaussian_pulse(p, seed=10 + i, center=5.0)
        rows = teuk.plunge_trajectory(p.a, 1e-2, 0.2, p.dt / 2, (2 * st + 6) * p.dt / 2, seed=i,
                                      r0_over_isco=1.3) if i != 2 else None
        specs.append((p, u0, rows))
    alone = [solve_alone(p, u0, rows, st) for (p, u0, rows), st in zip(specs, steps)]
    print(f"failed member integrator={integrator}", flush=True)
    solvers = []
    with teuk.Batch(integrator=integrator) as b:
        b.set_tiling(2, 2, 0, 16)
        for i, ((p, u0, rows), st) in enumerate(zip(specs, steps)):
            s = teuk.Solver(p)
            u = u0.copy()
            if i == 1:
                u[2, 8, 200] = np.nan             # p_re, row 8, column 200
            s.set_fields(u)
            if rows is not None:
                s.set_trajectory(rows, st)
            assert b.add(s)
            solvers.append(s)
        b.step(16)
        with pytest.raises(teuk.TeukError) as e:
            b.sync()
        assert e.value.code == teuk.TEUK_EDIVERGED
        assert [s.status() for s in solvers] == [teuk.TEUK_OK, teuk.TEUK_EDIVERGED, teuk.TEUK_OK]
        b.remove(solvers[1])
        b.step(max(steps))
        b.sync()
        for i in (0, 2):
            u, h, psi = alone[i]
            s = solvers[i]
            assert s.status() == teuk.TEUK_OK
            assert np.array_equal(s.get_fields(), u)
            assert np.array_equal(s.output()["h"], h)
            assert np.array_equal(s.psi(), psi)
    for s in solvers:
        s.close()


@pytest.mark.parametrize("split", [(40,), (32, 8), (7, 33)])
def test_source_switched_off_mid_run(split):
    """32 steps with the plunge source, then the source off and 40 more
    steps in the given calls: fields equal the oracle's (32 steps with the
    source, then 40 source-free steps) bit for bit."""
    on, off = 32, 40
    p = params(256, 32, on + off)
    u0 = workloads.gaussian_pulse(p, seed=5, center=20.0)
    rows = teuk.plunge_trajectory(p.a, 1e-2, 0.3, p.dt / 2, (2 * on + 6) * p.dt / 2, seed=5, r0_over_isco=1.3)
    _, r_c, th = teuk.grid(p)
    print(f"source off after {on} steps, then {split}", flush=True)
    with teuk.Solver(p, r_c, th) as s:
        s.set_fields(u0)
        s.set_trajectory(rows, on)
        s.step(on)
        s.sync()
        s.set_source(None)
        for n in split:
            s.step(n)
        s.sync()
        ug = s.get_fields()
    sched = teuk.trajectory_schedule(rows, on, p.dt)
    ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=1)
    uo = u0.copy()
    assert O.run(op, r_c, th, uo, on, src_steps=ost)[0] == 0
    assert O.run(op, r_c, th, uo, off)[0] == 0
    assert np.array_equal(ug, uo)
