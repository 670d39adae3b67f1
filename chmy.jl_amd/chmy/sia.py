"""Shallow-ice (SIA) time stepper and its adjoint (examples/sia.jl form).

Forward: the explicit loop of sia.jl:71-96 with the fused step kernel
(update_q! + q Dirichlet + update_H!) and the H Neumann batch; the time step
follows the reference's rule Δt = min(Δ²/max(D)/4.1, Δt_max) (sia.jl:76-78)
and the sequence is recorded so the adjoint differentiates the same discrete
map.  Gradient (build extension): J = ½‖H_N - H_obs‖² over the cells, d J /
d log A by hand-written adjoint steps in a native reverse sweep
(chmy_sia_reverse) with uniform checkpointing: `ncheck` stored states, each
segment recomputed once from its checkpoint and kept in HBM while its adjoint
steps run (the two-level revolve schedule); the recompute of the previous
segment shares launches with the adjoint steps.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .architectures import current_stream
from .boundary_conditions import Neumann, batch
from .fields import Field, VectorField, set_storage
from .function_field import FunctionField
from .grids import Center, Vertex, UniformGrid, spacing
from .kernels import _arr, maxabs


class SIA:
    def __init__(self, arch, grid, B, logA=None, *, As=0.0, npow=3, mb=(0.5, 0.3, 0.25), dt_max=math.inf):
        if grid.nd != 2:
            raise ValueError("SIA is 2-D")
        self.arch, self.grid = arch, grid
        self.As, self.npow = float(As), int(npow)
        self.mb = tuple(float(m) for m in mb)
        self.dt_max = float(dt_max)
        f = lambda loc=Center(): Field(arch, grid, loc, torch.float64)
        self.H, self.H2, self.B, self.D = f(), f(), f(), f()
        if isinstance(B, FunctionField):
            # the bed as a FunctionField (examples/sia.jl:30-36), evaluated on
            # the device over the whole storage
            B.materialize(arch, out=self.B)
        else:
            set_storage(self.B, B)
        self.A = None
        if logA is not None:
            self.A = f()
            self.set_logA(logA)
        self.grad = f() if logA is not None else None
        self.lam, self.lam2 = f(), f()
        self.work = [f(), f((Vertex(), Center())), f((Vertex(), Center())), f((Center(), Vertex())),
                     f((Center(), Vertex())), f(), f(), f()]
        self._work_arr = None
        self._bc = {}

    def set_logA(self, logA):
        """A = exp(logA) on the full storage (logA given on the full storage)."""
        self.logA = np.array(logA, dtype=np.float64)
        set_storage(self.A, np.exp(self.logA))

    # ------------------------------------------------------------------ forward
    def _neumann(self, H):
        bs = self._bc.get(id(H))
        if bs is None:
            bs = batch(self.grid, (H, Neumann()))
            self._bc[id(H)] = bs
        bs.apply()

    def reset(self, H0=None):
        self.H.data.zero_()
        if H0 is not None:
            set_storage(self.H, H0)
        self._neumann(self.H)

    def diffusivity_max(self, H=None):
        g = self.grid
        H = H or self.H
        _lib.call("chmy_sia_diffusivity", C.byref(g.desc()), C.byref(self.D.desc()), C.byref(H.desc()),
                  C.byref(self.B.desc()), C.byref(self.A.desc()) if self.A is not None else None, self.As,
                  self.npow, current_stream())
        return maxabs(self.D)

    def dt_rule(self, H=None):
        """Δt = min(min(Δ)²/max(D)/4.1, Δt_max) (sia.jl:76-78)."""
        dmax = self.diffusivity_max(H)
        dt_diff = min(spacing(self.grid)) ** 2 / dmax / 4.1 if dmax > 0 else math.inf
        return min(dt_diff, self.dt_max)

    def _args(self):
        A = C.byref(self.A.desc()) if self.A is not None else None
        return C.byref(self.B.desc()), A, self.As, self.npow, *self.mb

    def step(self, dt, H=None, H_out=None):
        """One step as the reference composes it: the fused update_q! + q
        Dirichlet + update_H! kernel, then the H Neumann batch."""
        g = self.grid
        H = H or self.H
        H_out = H_out or self.H2
        _lib.call("chmy_sia_step", C.byref(g.desc()), C.byref(H.desc()), C.byref(H_out.desc()),
                  *self._args(), float(dt), current_stream())
        self._neumann(H_out)

    def run(self, states, dts):
        """states[q] -> states[q+1] for the given Δt (one tiled launch per step,
        Neumann batch included)."""
        if len(states) < 2:
            return
        d = (C.c_double * (len(states) - 1))(*[float(x) for x in dts])
        _lib.call("chmy_sia_run", C.byref(self.grid.desc()), _arr(states), len(states), *self._args(), d,
                  current_stream())

    def _field(self):
        return Field(self.arch, self.grid, Center(), torch.float64)

    def forward(self, nsteps, dts=None, ncheck=0):
        """Run nsteps from the current H; returns the Δt sequence (the
        reference's adaptive rule when dts is None, evaluated on each state).
        With ncheck > 0 the state before every stride-th step, stride =
        ceil(nsteps/ncheck), is written straight into a checkpoint field."""
        stride = max(1, -(-nsteps // ncheck)) if ncheck else 0
        nck = -(-nsteps // stride) if stride else 0
        self.ckpt = getattr(self, "ckpt", [])[:nck]
        while len(self.ckpt) < nck:
            self.ckpt.append(self._field())
        self.stride = stride
        work = [self.H, self.H2]
        states = []
        for k in range(nsteps + 1):
            if stride and k % stride == 0 and k < nsteps:
                states.append(self.ckpt[k // stride])
            else:
                states.append(work[k % 2])
        if states[0] is not self.H:
            states[0].data.copy_(self.H.data)
        if dts is None:
            out = []
            for k in range(nsteps):
                out.append(self.dt_rule(states[k]))
                self.run(states[k:k + 2], out[-1:])
        else:
            out = [float(x) for x in dts[:nsteps]]
            self.run(states, out)
        if states[-1] is not self.H:
            self.H.swap(self.H2)
        return out

    def misfit(self, H_obs):
        d = self.H.interior() - H_obs
        return 0.5 * float(torch.sum(d * d))

    # ------------------------------------------------------------------ adjoint
    def _adjoint_step(self, Hk, dt, passes=1):
        """lam <- (d step / d H)^T lam at state Hk, grad += the logA part;
        passes=1 the tiled kernel, passes=5 the five gather kernels."""
        g = self.grid
        if passes == 5 and self._work_arr is None:
            self._work_arr = _arr(self.work)
        _lib.call("chmy_sia_step_adjoint", C.byref(g.desc()), C.byref(Hk.desc()), C.byref(self.lam.desc()),
                  C.byref(self.lam2.desc()), C.byref(self.grad.desc()) if self.A is not None else None,
                  C.byref(self.B.desc()), C.byref(self.A.desc()) if self.A is not None else None, self.As,
                  self.npow, *self.mb, float(dt), self._work_arr if passes == 5 else None, current_stream())
        self.lam.swap(self.lam2)

    def gradient(self, H0, H_obs, nsteps, dts, ncheck):
        """J(logA) and dJ/dlogA (cells) for the recorded Δt sequence: the
        checkpointed forward run, then the native reverse sweep."""
        self.reset(H0)
        dts = self.forward(nsteps, dts, ncheck)
        J = self.misfit(H_obs)
        self.grad.data.zero_()
        self.lam.data.zero_()
        self.lam2.data.zero_()
        self.lam.interior().copy_(self.H.interior() - H_obs)
        stride = self.stride
        self.seg = getattr(self, "seg", [])[:2 * (stride - 1)]
        while len(self.seg) < 2 * (stride - 1):
            self.seg.append(self._field())
        d = (C.c_double * nsteps)(*dts)
        slot = C.c_int(-1)
        seg = _arr(self.seg) if self.seg else None
        _lib.call("chmy_sia_reverse", C.byref(self.grid.desc()), _arr(self.ckpt), len(self.ckpt), stride, nsteps,
                  seg, C.byref(self.lam.desc()), C.byref(self.lam2.desc()),
                  C.byref(self.grad.desc()) if self.A is not None else None, *self._args(), d, C.byref(slot),
                  current_stream())
        if slot.value == 1:
            self.lam.swap(self.lam2)
        g = self.grad.interior().clone() if self.A is not None else None
        return J, g


def bed_bumps(grid, nbump=8, seed=4, hmin=500.0, hmax=1000.0, smin=20e3, smax=45e3, cmin=None, cmax=None):
    """Seeded sum of Gaussian bumps on the full storage of a Center field."""
    rng = np.random.default_rng(seed)
    L = grid.extent[0]
    cmin = cmin if cmin is not None else 0.2 * L
    cmax = cmax if cmax is not None else 0.8 * L
    xs = grid.coords(Center(), 0, first=-1, count=grid.dims[0] + 4) - grid.origin[0]
    ys = grid.coords(Center(), 1, first=-1, count=grid.dims[1] + 4) - grid.origin[1]
    X, Y = xs[None, :], ys[:, None]
    B = np.zeros((grid.dims[1] + 4, grid.dims[0] + 4))
    for _ in range(nbump):
        h = rng.uniform(hmin, hmax)
        cx, cy = rng.uniform(cmin, cmax, 2)
        s = rng.uniform(smin, smax)
        B += h * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / s ** 2)
    return B


def smooth_field(grid, sigma, seed, corr=8):
    """Seeded smooth random field (σ given) on the full storage (box-filtered noise)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((grid.dims[1] + 4, grid.dims[0] + 4))
    k = np.ones(corr) / corr
    for ax in (0, 1):
        a = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, a)
    a = (a - a.mean()) / a.std()
    return sigma * a
