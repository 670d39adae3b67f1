"""The physical time loop of examples/stokes_3d_inc_ve_T.jl (:152-184, 2D
stokes_2d_inc_ve_T.jl) on the fused PT iteration.

Per time step it: update_old! (T, τ → T_old, τ_old); dt = min(dt_diff,
dt_adv) with dt_adv = 0.1 min(Δ_D / max|V_D|) / N / 2.1 (∞ while V = 0);
η_ve = 1/(1/η + 1/(G dt)); then up to niter PT iterations, each the fused
update_stress! + update_velocity! with the V batch, followed from the second
time step on by update_thermal_flux!, update_thermal! and the T Neumann
batch (one launch into the second of two T buffers, then the batch on it:
the same values as the two launches in place; in the first step, with no
thermal update, the iterations between checks run in one native call; with
the thermal update, two iterations between checks run as one pass that also
keeps the first iteration's V for its thermal update, chmy_stokes_iter2_first);
every ncheck iterations the residual (∇V, r_V with Dirichlet-zeroed
boundary vertices, scaled by τsc, lz/psc) decides convergence, and from the
third step on ncheck = ceil(0.5 n_x).  With exact=True the states are
bitwise those of the reference's kernel sequence.
"""
from __future__ import annotations

import math

import torch

from .boundary_conditions import Neumann, batch
from .fields import Field, VectorField
from .grids import Center, spacing
from .kernel_launch import Launcher
from .kernels import maxabs, thermal_step, update_old
from .stokes import PTStokes


class StokesThermal:
    def __init__(self, arch, grid, *, eta=10.0, G=1.0, lam_rhoCp=None, rhog=None, normal_values=None,
                 re_m=2.3 * math.pi, r=0.5, exact=False, pairs=True):
        self.arch, self.grid = arch, grid
        self.eta, self.G = float(eta), float(G)
        L = grid.extent
        tau_sc = self.eta / self.G
        self.lam = float(lam_rhoCp) if lam_rhoCp is not None else 1e-4 * L[-1] ** 2 / tau_sc
        self.S = PTStokes(arch, grid, dtype=torch.float64, eta=eta, G=G, dt=1.0, normal_values=normal_values,
                          re_m=re_m, r=r, rhog=rhog, exact=exact)
        # T double-buffered: thermal_step reads one buffer and writes the other
        self._Tbuf = [Field(arch, grid, Center()), Field(arch, grid, Center())]
        self._Tbcs = [batch(grid, (f, Neumann())) for f in self._Tbuf]
        self._tp = 0
        self.T_old = Field(arch, grid, Center())
        self.qT = VectorField(arch, grid)
        self.launch = Launcher(arch, grid, sync=False)
        self.scales = dict(tau_sc=tau_sc, lz=L[-1], psc=self.G)
        # two PT iterations per pass inside the thermal PT loop, with the
        # first iteration's V kept in _V1 for its thermal update
        self.pairs = bool(pairs) and grid.nd in (2, 3)
        self._V1 = VectorField(arch, grid) if self.pairs else None

    @property
    def T(self):
        """The current temperature Field."""
        return self._Tbuf[self._tp]

    def apply_bc(self):
        self.S.apply_bc()
        self._Tbcs[self._tp].apply()
        self.S.sync_buffers()
        if self._V1 is not None:   # points no pass writes agree with V's
            for a, b in zip(self.S.V, self._V1):
                b.data.copy_(a.data)

    def dt_rule(self):
        g = self.grid
        d = spacing(g)
        dt_diff = min(d) ** 2 / self.lam / g.nd / 2.1
        vmax = [maxabs(self.S.V[D]) for D in range(g.nd)]
        dt_adv = 0.1 * min(d[D] / vmax[D] if vmax[D] > 0 else math.inf for D in range(g.nd)) / g.nd / 2.1
        return min(dt_diff, dt_adv)

    def time_step(self, it, niter, ncheck, eps):
        """One physical step (it is 1-based as in the reference); returns
        (PT iterations, last residual)."""
        S, g = self.S, self.grid
        self.launch(self.arch, g, (update_old, (self.T, S.tau, self.T_old, S.tau_old)))
        dt = self.dt_rule()
        S.dt = dt
        S.eta_ve_s = 1.0 / (1.0 / self.eta + 1.0 / (self.G * dt))
        err = None
        if it == 1:
            # no thermal update: the iterations between checks run in one
            # native call (in 2D two per pass), with the residual fields
            # written by the last one
            q = 0
            while q < niter:
                k = min(ncheck - q % ncheck, niter - q)
                check = (q + k) % ncheck == 0
                S.iterate(k, residual_last=check)
                q += k
                if check:
                    err = S.residual(**self.scales)
                    if all(e < eps for e in err):
                        return q, err
                    if not all(math.isfinite(e) for e in err):
                        raise FloatingPointError(f"simulation failed, err = {err}")
            return niter, err
        # both T buffers start equal, so points no launch writes agree
        self._Tbuf[1 - self._tp].data.copy_(self.T.data)
        q = 1
        while q <= niter:
            check = q % ncheck == 0
            if self.pairs and not check and q < niter and (q + 1) % ncheck != 0:
                # iterations q and q+1 in one pass; each thermal update reads
                # the V of its own iteration (Stokes does not read T)
                S.pass2_first(self._V1)
                self._thermal(self._V1, dt)
                self._thermal(S.V, dt)
                q += 2
                continue
            S.iteration(write_res=check)
            self._thermal(S.V, dt)
            if check:
                err = S.residual(**self.scales)
                if all(e < eps for e in err):
                    return q, err
                if not all(math.isfinite(e) for e in err):
                    raise FloatingPointError(f"simulation failed, err = {err}")
            q += 1
        return niter, err

    def _thermal(self, V, dt):
        """update_thermal_flux! + update_thermal! (one launch into the other T
        buffer) and the T Neumann batch on it."""
        g = self.grid
        src, dst = self._tp, 1 - self._tp
        self.launch(self.arch, g, (thermal_step, (self._Tbuf[src], self._Tbuf[dst], self.T_old, self.qT,
                                                  V, self.lam, dt, g)), bc=self._Tbcs[dst])
        self._tp = dst

    def run(self, nt=2, niter=None, ncheck=None, eps=1e-6):
        """The reference's defaults: niter = 50 n_x, ncheck = 2 n_x, then
        ceil(0.5 n_x) from the third step on."""
        nx = self.grid.dims[0]
        niter = niter or 50 * nx
        ncheck = ncheck or 2 * nx
        out = []
        for it in range(1, nt + 1):
            if it > 2:
                ncheck = math.ceil(0.5 * nx)
            out.append(self.time_step(it, niter, ncheck, eps))
        return out
