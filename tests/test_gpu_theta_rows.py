"""Stage-kernel wavefronts of 2, 4 and 8 theta rows against the oracle, bit
for bit.  The first and last theta chunk of every rho strip compute 2, 3 or
5 rows next to the boundary rows they fill and the chunks between 2, 4 or 8
rows (M = 16 with 8-row wavefronts: edge chunks only; M = 24 and 40: chunk
counts that 2- and 4-wave blocks do not divide, so surplus wavefronts), with
a source patch that covers the whole theta range and so crosses every chunk
boundary, for both integrators."""
import numpy as np
import pytest

from bh_ringdown_amd import teuk
from oracle import oracle as O

from test_gpu_source import assert_same_bits, inspiral_rows

pytestmark = pytest.mark.gpu

GRIDS = [(256, 16), (256, 24), (128, 40), (512, 64)]


def run_both(p, steps, rows, wpb, theta_rows):
    _, r_c, th = teuk.grid(p)
    with teuk.Solver(p, r_c, th) as s:
        s.set_tiling(wpb, 2, theta_rows)
        s.set_trajectory(rows, steps)
        s.step(steps)
        s.sync()
        ug, out, psig = s.get_fields(), s.output(), s.psi()
    sched = teuk.trajectory_schedule(rows, steps, p.dt)
    op = O.make_params(p.N, p.M, p.xmin, p.xmax, p.a, p.m, dt=p.dt, integrator=p.integrator,
                       j_extract=p.j_extract)
    ost = [O.SrcStep(*[getattr(x, f) for f, _ in teuk.SourceStep._fields_]) for x in sched]
    uo = np.zeros((4, p.M, p.N))
    rc, psio, ho, _ = O.run(op, r_c, th, uo, steps, src_steps=ost)
    assert rc == 0
    return ug, uo, out, psig, psio, ho


@pytest.mark.parametrize("integrator", [1, 0])
@pytest.mark.parametrize("N,M", GRIDS)
@pytest.mark.parametrize("theta_rows,wpb", [(2, 1), (2, 4), (4, 1), (4, 2), (8, 1), (8, 4)])
def test_theta_rows_bitwise_against_oracle(theta_rows, wpb, N, M, integrator):
    steps = 60
    p = teuk.default_params(N, M, -40.0, 60.0, 0.7, 2, dt_over_dx=0.5)
    p.T = steps * p.dt * (1 + 1e-12)
    p.integrator = integrator
    p.j_extract = N - 30
    rows = inspiral_rows(2 * steps + 4, p.dt, r0=8.0)
    print(f"theta rows {theta_rows} wpb {wpb} grid {N}x{M} integ {integrator} steps {steps}", flush=True)
    ug, uo, out, psig, psio, ho = run_both(p, steps, rows, wpb, theta_rows)
    assert np.any(uo != 0.0)
    assert_same_bits(ug, uo)
    assert_same_bits(psig, psio)
    assert np.allclose(out["h"], ho, rtol=1e-8, atol=1e-8 * np.abs(ho).max())


def test_launch_shapes_rejected_outside_their_ranges():
    """Theta rows other than 2, 4, 8 and waves per SIMD other than 2-4 are
    refused, for solvers and batches; the staggered scheme has no stage
    kernel and refuses both setters."""
    p = teuk.default_params(256, 16, -40.0, 60.0, 0.7, 2, dt_over_dx=0.5)
    p.T = 5 * p.dt * (1 + 1e-12)
    with teuk.Solver(p) as s:
        for bad in (0, 1, 3, 16):
            with pytest.raises(teuk.TeukError):
                s.set_tiling(1, 2, bad)
        for wpe in (1, 5):
            with pytest.raises(teuk.TeukError):
                s.set_tiling(1, wpe)
        s.set_tiling(1, 3, 4)      # in range: accepted
        s.step(3)
        s.sync()
    with teuk.Batch(integrator=1) as b:
        for bad in (0, 3, 6):
            with pytest.raises(teuk.TeukError):
                b.set_tiling(1, 2, 0, 32, bad)
        with pytest.raises(teuk.TeukError):
            b.set_tiling(1, 5, 0, 32)
    q = teuk.default_params(256, 16, -40.0, 60.0, 0.7, 2, dt_over_dx=0.5)
    q.integrator = 2
    with teuk.Solver(q) as s:
        with pytest.raises(teuk.TeukError):
            s.set_tiling(1, 2)
        with pytest.raises(teuk.TeukError):
            s.set_tiling(1, 2, 2)
