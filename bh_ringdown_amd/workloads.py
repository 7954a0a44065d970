"""Named workloads of the 2+1D Teukolsky solver and their synthetic inputs.

Each workload fixes the grid (rho domain = the reference's [Xmin, Xmax]),
spin, mode, step count and initial data.  Sizes are the ones the solver is
measured and checked on.
"""
import math

import numpy as np

from . import teuk


def j_at(p, rho):
    """Extraction column closest to rho (the reference extracts at N-1)."""
    return int(round((rho - p.xmin) / p.dx))


def gaussian_pulse(p, seed, center=50.0, width=5.0):
    """Complex Gaussian Psi = A e^{i phi0} exp(-(rho-c)^2/(2 w^2)) sin^4(theta),
    Pi = 0, with (A, phi0) drawn from the seed (A ~ U[0.5, 1.5), phi0 ~ U[0, 2 pi))."""
    rng = np.random.default_rng(seed)
    amp = 0.5 + rng.random()
    phase = 2.0 * math.pi * rng.random()
    x = p.xmin + np.arange(p.N) * p.dx
    th = np.arange(p.M) * (teuk.PI15 / p.M)
    prof = np.exp(-((x - center) ** 2) / (2.0 * width * width))
    ang = np.sin(th) ** 4
    g = amp * np.outer(ang, prof)
    u = np.zeros((4, p.M, p.N))
    u[0] = g * math.cos(phase)
    u[1] = g * math.sin(phase)
    return u


def plain_1024x32(steps=2000, integrator=1, Fr=1):
    """Host-oracle parity case (the plain method): s=-2, m=2, a=0.5; N=1024
    on rho in [-100,300], M=32; classic RK4 (integrator=0: the reference's
    SSP-RK3); dt = 0.5 dx; no source; Gaussian initial Psi; Psi at rho=250."""
    p = teuk.default_params(1024, 32, -100.0, 300.0, 0.5, 2, dt_over_dx=0.5)
    p.T = steps * p.dt * (1.0 + 1e-12)
    p.Fr = Fr
    p.integrator = integrator
    p.j_extract = j_at(p, 250.0)
    return p


def single_16384x64(T=1500.0, integrator=0, Fr=2):
    """Single full solve: s=-2, m=2, a=0.9; N=16384 on rho in [-400,1200],
    M=64; dt = 0.5 dx to t=1500 (30720 steps); Psi at rho=1000."""
    p = teuk.default_params(16384, 64, -400.0, 1200.0, 0.9, 2, dt_over_dx=0.5, T=T)
    p.Fr = Fr
    p.integrator = integrator
    p.j_extract = j_at(p, 1000.0)
    return p


HIGH_SPIN_GRIDS = ((16384, 32), (32768, 64), (65536, 128))


def high_spin(N, M, T=2000.0, integrator=1, Fr=2):
    """High-spin resolution ladder: s=-2, m=2, a=0.99 on rho in [-600,1600]
    (layer from rho=808), dt = 0.5 drho to t=2000; one of HIGH_SPIN_GRIDS
    (16384x32: 29789 steps; 32768x64: 59578; 65536x128: 119156, 8.4M points).
    Psi at rho=1400."""
    p = teuk.default_params(N, M, -600.0, 1600.0, 0.99, 2, dt_over_dx=0.5, T=T)
    p.Fr = Fr
    p.integrator = integrator
    p.j_extract = j_at(p, 1400.0)
    return p


def plunge_rows(p, nsteps, eta=1e-3, iota_deg=30.0, seed=0, r0_over_isco=1.02):
    """Seeded transition-plus-plunge .traj rows for spin p.a, sampled at dt/2
    and long enough for nsteps of the reference's two-rows-per-step reading."""
    t_final = (2 * nsteps + 4) * (p.dt / 2)
    return teuk.plunge_trajectory(p.a, eta, math.radians(iota_deg), p.dt / 2, t_final, seed=seed,
                                  r0_over_isco=r0_over_isco)


def grid_for_spin(a):
    """Grid size by spin for the batched set: higher spins need more rho
    resolution near the horizon (and theta resolution at the top)."""
    if a < 0.4:
        return 4096, 32
    if a < 0.8:
        return 8192, 32
    return 16384, 64


def batch_72(seed=0, t_final=1200.0, integrator=1, eta=1e-3, iota_deg=30.0, grid=None):
    """Batched uneven solves: m in -4..4 crossed with 8 spins a ~ U[0, 0.99)
    drawn from the seed, each spin's grid from grid_for_spin() on rho in
    [-400, 1200], dt = 0.5 drho, to t_final, all with the same inclined
    plunge (mass ratio eta, inclination iota) in each spin's geometry.
    grid=(N, M) gives every solve that grid instead.  Returns
    [(params, trajectory rows)] (72 entries)."""
    rng = np.random.default_rng(seed)
    spins = rng.uniform(0.0, 0.99, size=8)
    out = []
    for a in spins:
        N, M = grid if grid else grid_for_spin(a)
        rows = None
        for m in range(-4, 5):
            p = teuk.default_params(N, M, -400.0, 1200.0, float(a), m, dt_over_dx=0.5, T=t_final)
            p.integrator = integrator
            p.Fr = 2
            if rows is None:
                rows = plunge_rows(p, p.nsteps(), eta=eta, iota_deg=iota_deg, seed=seed)
            out.append((p, rows))
    return out
