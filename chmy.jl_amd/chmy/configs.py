"""Synthetic setups of the benchmark workloads (seeded, analytic).

- stokes3d_inclusion: 3D VE Stokes, [-0.5,0.5]³, background η = 1, sphere of
  radius 0.1 with η = 1e-3, times exp(-T), T = 0.1·exp(-|x|²/0.05²), G = 1,
  dt from the reference's diffusion limit dt = Δ²/λ/N/2.1 with λ = 1e-4
  (examples/stokes_3d_inc_ve_T.jl:93,157), pure-shear free-slip walls ±1.
- stokes2d_inclusions: 2D VE Stokes, [-0.5,0.5]², 32 circular inclusions at
  seeded positions (centres U[-0.5,0.5]², radii U[0.01,0.04], log10 η
  U[-3,3]; later inclusions overwrite earlier ones), background η = 1,
  G = 1, same dt rule, pure shear ±1.
- sia_ice_cap: 2D SIA, [0, 250 km]², 8 seeded Gaussian bed bumps (heights
  U(0, 1] km, centres U[0.3L, 0.7L]², widths U[30, 60] km), H0 = 0, mass
  balance min(0.01 (S - 1500 m), 2 m/a) (ELA 1.5 km), n = 3, log A = log(1e-4)
  plus, for the observation, a smooth seeded perturbation of σ = 0.3; fixed
  Δt = 0.25 a.
η is returned on the full padded storage (logical -1..n+2 per dim) evaluated
at the fma cell-centre coordinates.
"""
from __future__ import annotations

import numpy as np

from .grids import Center, UniformGrid, spacing


def dt_diffusion(grid, lam=1e-4):
    return min(spacing(grid)) ** 2 / lam / grid.nd / 2.1


def stokes3d_inclusion(arch, n=255, *, eta0=1.0, eta_incl=1e-3, radius=0.1, T_amp=0.1, T_width=0.05):
    g = UniformGrid(arch, origin=(-0.5,) * 3, extent=(1.0,) * 3, dims=(n,) * 3)
    xs = [g.coords(Center(), d, first=-1, count=g.dims[d] + 4) for d in range(3)]
    X = xs[0][None, None, :]
    Y = xs[1][None, :, None]
    Z = xs[2][:, None, None]
    r2 = X ** 2 + Y ** 2 + Z ** 2
    T = T_amp * np.exp(-r2 / T_width ** 2)
    eta = np.where(r2 < radius ** 2, eta_incl, eta0) * np.exp(-T)
    return g, eta, dt_diffusion(g)


def stokes2d_inclusions(arch, n=2047, *, count=32, seed=2047, rmin=0.01, rmax=0.04, lmin=-3.0, lmax=3.0):
    g = UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.0), dims=(n, n))
    rng = np.random.default_rng(seed)
    cx = rng.uniform(-0.5, 0.5, count)
    cy = rng.uniform(-0.5, 0.5, count)
    rad = rng.uniform(rmin, rmax, count)
    le = rng.uniform(lmin, lmax, count)
    xs = g.coords(Center(), 0, first=-1, count=n + 4)
    ys = g.coords(Center(), 1, first=-1, count=n + 4)
    X, Y = xs[None, :], ys[:, None]
    eta = np.ones((n + 4, n + 4))
    for q in range(count):
        eta = np.where((X - cx[q]) ** 2 + (Y - cy[q]) ** 2 < rad[q] ** 2, 10.0 ** le[q], eta)
    return g, eta, dt_diffusion(g)


def sia_ice_cap(arch, n=512, *, seed=0, sigma=0.3, dt=0.25, logA0=float(np.log(1e-4))):
    """Grid, bed B, uniform log A, perturbed log A (for H_obs), mass-balance
    triple and the fixed Δt of the SIA gradient workload (full storage arrays)."""
    from .sia import bed_bumps, smooth_field
    g = UniformGrid(arch, origin=(0.0, 0.0), extent=(250e3, 250e3), dims=(n, n))
    L = g.extent[0]
    B = bed_bumps(g, nbump=8, seed=seed, hmin=0.0, hmax=1000.0, smin=30e3, smax=60e3, cmin=0.3 * L, cmax=0.7 * L)
    la = np.full((n + 4, n + 4), logA0)
    return g, B, la, la + smooth_field(g, sigma, seed=seed + 1, corr=16), (0.01, 1500.0, 2.0), dt
