"""teuk_solve_noalloc (the entry point that allocates nothing) against the
same solve through the allocating entry points (teuk_create, set_fields,
set_source, step, sync, get_fields, read_output), bit for bit, on two calls
in a row -- the second smaller than the first -- and its refusal of a solve
larger than its fixed scratch."""
import numpy as np
import pytest

from bh_ringdown_amd import teuk, workloads

pytestmark = pytest.mark.gpu


def case(N, M, a, m, steps, integrator, Fr, source, seed):
    p = teuk.default_params(N, M, -60.0, 100.0, a, m, dt_over_dx=0.5)
    p.T = steps * p.dt * (1 + 1e-12)
    p.integrator, p.Fr = integrator, Fr
    _, r_c, th = teuk.grid(p)
    u0 = workloads.gaussian_pulse(p, seed=seed, center=20.0)
    sched = None
    if source:
        rows = teuk.plunge_trajectory(a, 1e-2, 0.3, p.dt / 2, (2 * steps + 6) * p.dt / 2, seed=seed,
                                      r0_over_isco=1.3)
        sched = teuk.trajectory_schedule(rows, steps, p.dt)
    return p, r_c, th, u0, sched


def allocating(p, r_c, th, u0, sched, steps):
    with teuk.Solver(p, r_c, th) as s:
        s.set_fields(u0)
        if sched is not None:
            s.set_source(sched)
        s.step(steps)
        s.sync()
        return s.get_fields(), s.output()


def test_noalloc_equals_allocating_path_twice_then_smaller():
    cases = [(1024, 64, 0.7, 2, 100, 1, 3, True, 1),      # RK4 with the plunge source, 3 graph-sized chunks + 4
             (256, 16, 0.4, -1, 45, 0, 2, False, 2)]      # smaller, SSP-RK3, no source
    for N, M, a, m, steps, integ, Fr, src, seed in cases:
        print(f"noalloc N={N} M={M} integrator={integ} steps={steps} source={src}", flush=True)
        p, r_c, th, u0, sched = case(N, M, a, m, steps, integ, Fr, src, seed)
        rc, u, out = teuk.solve_noalloc(p, r_c, th, steps, u0=u0, steps=sched)
        assert rc == teuk.TEUK_OK
        ua, oa = allocating(p, r_c, th, u0, sched, steps)
        assert np.array_equal(u, ua)
        for key in ("k", "write", "t", "dedt", "h"):
            assert np.array_equal(out[key], oa[key]), key
        assert len(out["k"]) == len(range(0, steps, Fr))


def test_noalloc_refuses_a_solve_beyond_its_scratch():
    p = teuk.default_params(16384, 64, -400.0, 1200.0, 0.9, 2, dt_over_dx=0.5, T=10.0)
    p.integrator = 1
    _, r_c, th = teuk.grid(p)
    rc, u, out = teuk.solve_noalloc(p, r_c, th, 100)
    assert rc == teuk.TEUK_ETOOBIG
    assert len(out["k"]) == 0 and not u.any()
    assert "scratch" in teuk.strerror(rc)
