"""Pin the oracle (and libteuk's host set-up) to the reference's own C code.

oracle/_ref/libref_host.so is TD_GPU_code/HighOrder/main.c + proj.c compiled
as they lie (parm.h: N=16000, M=256, rho in [-50, 50]); it is built from
/root/reference by oracle/Makefile.  Skipped where it has not been built.
"""
import numpy as np
import pytest

from oracle import oracle as O

ref = pytest.mark.skipif(O.ref_host() is None, reason="oracle/_ref not built")


@ref
@pytest.mark.parametrize("a", [1e-8, 0.3, 0.5, 0.9, 0.99])
def test_grid_bitwise_vs_reference(a):
    xr, rr, thr = O.ref_grid(1.0, a)
    p = O.make_params(O.REF_N, O.REF_M, O.REF_XMIN, O.REF_XMAX, a, 2)
    x, r, th = O.grid(p)
    assert np.array_equal(x, xr)
    assert np.array_equal(th, thr)
    fin = np.isfinite(rr)
    assert np.array_equal(r[fin], rr[fin])


@ref
def test_libteuk_host_grid_bitwise_vs_reference():
    from bh_ringdown_amd import teuk
    for a in (1e-8, 0.5, 0.9):
        p = teuk.default_params(O.REF_N, O.REF_M, O.REF_XMIN, O.REF_XMAX, a, 2)
        x, r, th = teuk.grid(p)
        xr, rr, thr = O.ref_grid(1.0, a)
        assert np.array_equal(r, rr) and np.array_equal(th, thr) and np.array_equal(x, xr)


@ref
@pytest.mark.parametrize("m", [-4, -2, -1, 0, 1, 2, 3, 5])
def test_project_bitwise_vs_reference(m):
    rng = np.random.default_rng(abs(m) + 10)
    qq = rng.standard_normal(O.REF_M)
    for l in range(abs(m), abs(m) + 7):
        assert O.project(qq, O.REF_M, l, m) == O.ref_project(qq, l, m)


@ref
def test_swsh_bitwise_vs_reference():
    l = O.ref_host()
    for ell in range(0, 9):
        for m in range(-ell, ell + 1):
            for th in (0.0, 0.3, 1.2, 2.9, np.pi):
                assert O.swsh(ell, m, th) == l.swsh(float(ell), float(m), th)
