"""Solvers driven from concurrent host threads: device state is per handle
(no static device globals, unlike kern.cu:10-36), so modes may be solved at
the same time from different threads.  Each thread's result must be bit for
bit the one of a serial run."""
import threading

import numpy as np
import pytest

from bh_ringdown_amd import teuk, workloads

pytestmark = pytest.mark.gpu


def case(m, integrator, seed):
    p = teuk.default_params(512, 32, -40.0, 60.0, 0.6, m, dt_over_dx=0.5)
    steps = 90
    p.T = steps * p.dt * (1 + 1e-12)
    p.integrator = integrator
    p.Fr = 3
    u0 = workloads.gaussian_pulse(p, seed=seed, center=10.0, width=4.0)
    rows = teuk.plunge_trajectory(p.a, 1e-2, 0.4, p.dt / 2, (2 * steps + 6) * p.dt / 2, seed=seed,
                                  r0_over_isco=1.3)
    return p, u0, rows, steps


def solve(p, u0, rows, steps):
    with teuk.Solver(p) as s:
        s.set_fields(u0)
        s.set_trajectory(rows, steps)
        for _ in range(3):              # several calls, interleaving with the other threads
            s.step(steps // 3)
        s.sync()
        return s.get_fields(), s.output()["h"], s.psi()


def test_solvers_on_concurrent_host_threads():
    cases = [case(m, integ, seed) for m, integ, seed in ((2, 1, 1), (-3, 1, 2), (1, 0, 3), (4, 2, 4))]
    serial = [solve(*c) for c in cases]
    got = [None] * len(cases)
    errors = []

    def work(i):
        try:
            got[i] = solve(*cases[i])
        except Exception as e:  # reported below, on the main thread
            errors.append((i, e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert all(not t.is_alive() for t in threads)
    for (u, h, psi), (u2, h2, psi2) in zip(serial, got):
        assert np.array_equal(u, u2)
        assert np.array_equal(h, h2)
        assert np.array_equal(psi, psi2)
