"""Gradient of a steady PT Stokes solve with respect to log viscosity
(build extension; the reference has no adjoint).

J(log η) = ½ Σ over the top cell layer (last-dimension index n, all points
of the tangential velocity arrays there: Vx, Vy in 3D, Vx in 2D) of
(V - V_obs)², with V the converged solution of
the viscoelastic Stokes problem of examples/stokes_3d_inc_ve_T.jl
(τ_old = 0, free-slip walls with the given normal velocities) and η a Center
field on the full storage.  The steady operator is symmetric up to the sign
of the pressure, so the adjoint is the same PT problem with forcing
f_V = -∂J/∂V, homogeneous boundary values and τ_old = 0; its stresses τ_a and
the forward stresses τ give dJ/dlog η in one pass (chmy_stokes_grad_eta).
Both solves use the fused PT iteration, in f64 or as FP32 sweeps with FP64
defect correction (MixedPTStokes).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .architectures import current_stream
from .fields import Field, VectorField
from .grids import Center
from .kernels import _arr
from .stokes import MixedPTStokes, PTStokes, pure_shear_initial


class StokesAdjoint:
    def __init__(self, arch, grid, logeta, *, G=1.0, dt=1.0, normal_values=None, mixed=False, inner=200,
                 exact=False):
        self.arch, self.grid, self.mixed = arch, grid, bool(mixed)
        nd = grid.nd
        eta = np.exp(np.asarray(logeta, dtype=np.float64))
        zero_bc = {D: (0.0, 0.0) for D in range(nd)}
        if self.mixed:
            mk = lambda nv: MixedPTStokes(arch, grid, eta_field=eta, G=G, dt=dt, normal_values=nv, inner=inner,
                                          exact=exact)
        else:
            mk = lambda nv: PTStokes(arch, grid, dtype=torch.float64, eta_field=eta, G=G, dt=dt, normal_values=nv,
                                     exact=exact)
        self.fwd = mk(normal_values)
        self.adj = mk(zero_bc)
        self.f_adj = VectorField(arch, grid, torch.float64)
        self._state(self.adj).f_V = self.f_adj
        if not self.mixed:   # the fused iteration takes f_V and f_P together
            self.adj.f_P = Field(arch, grid, Center(), torch.float64)
        self.grad = Field(arch, grid, Center(), torch.float64)
        self.eta = self._state(self.fwd).eta

    @staticmethod
    def _state(S):
        """The f64 PTStokes holding the solution (the master of a mixed solve)."""
        return S.x if isinstance(S, MixedPTStokes) else S

    def set_logeta(self, logeta):
        eta = np.exp(np.asarray(logeta, dtype=np.float64))
        self.fwd.set_eta(eta)
        self.adj.set_eta(eta)

    def _solve(self, S, eps, max_iter, V_init):
        S.reset_state(V_init)
        if isinstance(S, MixedPTStokes):
            return S.solve(eps=eps, max_cycles=max_iter // S.inner + 1)
        return S.solve(eps=eps, max_iter=max_iter)

    def solve_forward(self, eps=1e-8, max_iter=400000, V_init=pure_shear_initial):
        return self._solve(self.fwd, eps, max_iter, V_init)

    def top(self, f):
        """View of the top cell layer (last-dimension index n) of a tangential
        velocity array."""
        return f.interior()[..., -1]

    def misfit(self, V_obs):
        """J and ∂J/∂V (Vx, Vy on the top cell layer); V_obs: the top-layer
        arrays of an observed state, as returned by observe()."""
        V = self._state(self.fwd).V
        J = 0.0
        for f in self.f_adj:
            f.data.zero_()
        for D in range(self.grid.nd - 1):
            d = self.top(V[D]) - V_obs[D]
            J += 0.5 * float(torch.sum(d * d))
            self.top(self.f_adj[D]).copy_(-d)   # adjoint forcing -dJ/dV
        return J

    def observe(self):
        V = self._state(self.fwd).V
        return [self.top(V[D]).clone() for D in range(self.grid.nd - 1)]

    def solve_adjoint(self, eps=1e-8, max_iter=400000):
        return self._solve(self.adj, eps, max_iter, None)

    def gradient(self):
        """dJ/dlog η at centres 0..n+1 (zero elsewhere) after both solves."""
        g = self.grid
        self.grad.data.zero_()
        _lib.call("chmy_stokes_grad_eta", C.byref(g.desc()), _arr(self._state(self.fwd).tau),
                  _arr(self._state(self.adj).tau), C.byref(self.eta.desc()), C.byref(self.grad.desc()), 1,
                  current_stream())
        return self.grad

    def J_and_gradient(self, V_obs, eps=1e-8, max_iter=400000):
        self.solve_forward(eps, max_iter)
        J = self.misfit(V_obs)
        self.solve_adjoint(eps, max_iter)
        return J, self.gradient()


def smooth_log_field(grid, sigma, seed, corr=4):
    """Seeded smooth random field on the full Center storage (box-filtered
    Gaussian noise, zero mean, standard deviation sigma)."""
    rng = np.random.default_rng(seed)
    shape = tuple(n + 4 for n in reversed(grid.dims))
    a = rng.standard_normal(shape)
    k = np.ones(corr) / corr
    for ax in range(a.ndim):
        a = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, a)
    a = (a - a.mean()) / a.std()
    return sigma * a
