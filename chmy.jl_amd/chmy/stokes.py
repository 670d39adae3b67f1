"""Pseudo-transient viscoelastic Stokes solver on the MI355X library.

The time/PT loop of examples/stokes_3d_inc_ve_T.jl:152-184 (2D:
stokes_2d_inc_ve_T.jl) for one physical time step, driven from the host:
each PT iteration is one fused stress+velocity kernel (chmy_stokes_iter,
double-buffered) plus the free-slip V batch; every `ncheck` iterations the
kernel also writes ∇V and r_V, the boundary residual vertices are zeroed
(Dirichlet on r_V, :171) and the scaled L∞ norms are reduced on the device
(:172-175).

Variable viscosity (a build extension; the reference uses a scalar η): η is a
Center field given on its full padded storage; η_ml (its max over the 3^N
cell neighbourhood) sets the velocity pseudo-step.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .architectures import current_stream
from .boundary_conditions import Dirichlet, Neumann, batch
from .fields import Field, TensorField, VectorField, set_storage
from .function_field import as_field
from .grids import Center, UniformGrid, spacing
from .kernels import _arr, _box, maxloc, HipKernel
from .kernel_launch import Launcher


ITER_RESIDUAL, ITER_EXACT, ITER_PAIRS, ITER_SINGLE = 1, 2, 4, 8


def pt_params(grid, re_m=2.3 * math.pi, r=0.5):
    """PT parameters (stokes_3d_inc_ve_T.jl:105-112)."""
    lt = min(grid.extent) / re_m
    vdt = min(spacing(grid)) / math.sqrt(grid.nd * 1.1)
    th = lt * (r + 4 / 3) / vdt
    return dict(dtau_r=1.0 / (th + 1.0), nudtau=vdt * lt, dtau_Pr=r / th)


def free_slip_batch(grid, V, normal_values):
    """V.x => (x=Dirichlet(vl, vr), y=Neumann(), z=Neumann()), ... (stokes_3d_inc_ve_T.jl:130-132)
    with per-side normal velocities normal_values[D] = (left, right)."""
    names = grid.axes_names()
    pairs = []
    for D in range(grid.nd):
        spec = {n: Neumann() for n in names}
        vl, vr = normal_values[D]
        spec[names[D]] = (Dirichlet(vl), Dirichlet(vr))
        pairs.append((V[D], spec))
    return batch(grid, *pairs)


class PTStokes:
    """Fields and loop of one PT solve.  dtype float64 or float32."""

    def __init__(self, arch, grid, *, dtype=torch.float64, eta=None, eta_field=None, G=1.0, dt=None,
                 normal_values=None, re_m=2.3 * math.pi, r=0.5, rhog=None, exact=False):
        self.arch, self.grid, self.dtype = arch, grid, dtype
        nd = grid.nd
        self.nd = nd
        mk = lambda: (VectorField(arch, grid, dtype), TensorField(arch, grid, dtype), Field(arch, grid, Center(), dtype))
        self.V, self.tau, self.Pr = mk()
        self.V2, self.tau2, self.Pr2 = mk()
        self.tau_old = TensorField(arch, grid, dtype)
        self.divV = Field(arch, grid, Center(), dtype)
        self.r_V = VectorField(arch, grid, dtype)
        # a FunctionField ρg (examples/stokes_2d_inc_ve_T.jl:107) is evaluated
        # on the device into a field of the solver's dtype
        self.rhog = as_field(rhog, arch, dtype)
        self.f_V = None          # defect-correction forcing (set by MixedPTStokes)
        self.f_P = None
        self.G = float(G)
        self.dt = float(dt) if dt is not None else 1.0
        self.eta_s = float(eta) if eta is not None else 1.0
        self.eta_ve_s = 1.0 / (1.0 / self.eta_s + 1.0 / (self.G * self.dt))
        self.eta = None
        self.eta_ml = None
        if eta_field is not None:
            self.eta = Field(arch, grid, Center(), dtype)
            set_storage(self.eta, eta_field)
            self.eta_ml = Field(arch, grid, Center(), dtype)
            Launcher(arch, grid)(arch, grid, (maxloc, (self.eta_ml, self.eta)))
        self.p = pt_params(grid, re_m, r)
        nv = normal_values or {D: (0.0, 0.0) for D in range(nd)}
        self.normal_values = nv
        self._bcs = {}
        self._rbc = {}
        self.iters = 0
        # exact=True: reference expression order (bitwise equal to the two
        # reference kernels); False: the algebraically equal fast form
        self.exact = bool(exact)

    # ------------------------------------------------------------------ state
    def set_eta(self, eta_field):
        """Replace the viscosity field (full storage array) and its maxloc."""
        set_storage(self.eta, eta_field)
        Launcher(self.arch, self.grid)(self.arch, self.grid, (maxloc, (self.eta_ml, self.eta)))

    def reset_state(self, V_init=None):
        """Zero (V, τ, Pr) in both buffer sets, optionally start V from
        V_init(grid, V), then apply the V batch and sync the buffers."""
        for f in list(self.V) + list(self.tau) + [self.Pr] + list(self.V2) + list(self.tau2) + [self.Pr2]:
            f.data.zero_()
        if V_init is not None:
            V_init(self.grid, self.V)
        self.apply_bc()
        self.sync_buffers()

    def sync_buffers(self):
        """Make the output buffers equal to the inputs (also outside the box)."""
        for a, b in zip(list(self.V) + list(self.tau) + [self.Pr], list(self.V2) + list(self.tau2) + [self.Pr2]):
            b.data.copy_(a.data)

    def apply_bc(self, V=None):
        """Free-slip V batch on the given vector field (default: the current state).
        Batch sets are cached per Field object; they track storage swaps."""
        V = V if V is not None else self.V
        key = id(V[0])
        bs = self._bcs.get(key)
        if bs is None:
            bs = free_slip_batch(self.grid, V, self.normal_values)
            self._bcs[key] = bs
        bs.apply()

    def _iter_call(self, write_res):
        g = self.grid
        _lib.call("chmy_stokes_iter", C.byref(g.desc()), _arr(self.V), _arr(self.tau), C.byref(self.Pr.desc()),
                  _arr(self.V2), _arr(self.tau2), C.byref(self.Pr2.desc()), _arr(self.tau_old),
                  C.byref(self.eta.desc()) if self.eta is not None else None,
                  C.byref(self.eta_ml.desc()) if self.eta_ml is not None else None,
                  C.byref(self.rhog.desc()) if self.rhog is not None else None,
                  _arr(self.f_V) if self.f_V is not None else None,
                  C.byref(self.f_P.desc()) if self.f_P is not None else None,
                  C.byref(self.divV.desc()), _arr(self.r_V),
                  (ITER_RESIDUAL if write_res else 0) | (ITER_EXACT if self.exact else 0), self.eta_s, self.eta_ve_s, self.G,
                  self.dt, self.p["dtau_Pr"], self.p["dtau_r"], self.p["nudtau"], current_stream())

    def _swap(self):
        for a, b in zip(list(self.V) + list(self.tau) + [self.Pr], list(self.V2) + list(self.tau2) + [self.Pr2]):
            a.swap(b)

    def iteration(self, write_res=False):
        """One PT iteration: fused kernel into the output buffers, V batch on
        them, then swap so that (V, tau, Pr) hold the new state."""
        self._iter_call(write_res)
        self.apply_bc(self.V2)
        self._swap()
        self.iters += 1

    def _vbatch(self):
        """The free-slip V batch as (counts, component index, condition) arrays
        for chmy_stokes_iterate, built once."""
        if getattr(self, "_vb", None) is None:
            bs = free_slip_batch(self.grid, self.V, self.normal_values)
            counts = (C.c_int32 * 6)()
            comp, conds, keep = [], [], []
            for d in range(self.nd):
                for s in range(2):
                    b = bs[d][s]
                    counts[2 * d + s] = len(b.fields)
                    for f, cond in zip(b.fields, b.conditions):
                        comp.append(next(c for c in range(self.nd) if self.V[c] is f))
                        conds.append(cond.desc(keep))
            n = max(1, len(comp))
            self._vb = (counts, (C.c_int32 * n)(*comp), (_lib.ChmyBC * n)(*conds), keep)
        return self._vb

    def iterate(self, niter, residual_last=False, pairs=True):
        """niter PT iterations in one native call (fused kernel + V batch,
        buffers alternating); (V, tau, Pr) hold the result afterwards.  With
        the free-slip batch they run in pairs (one pass per two iterations,
        bitwise the same results); pairs=False keeps single iterations."""
        if niter <= 0:
            return
        g = self.grid
        counts, comp, conds, _keep = self._vbatch()
        base = (ITER_EXACT if self.exact else 0) | (0 if pairs else ITER_SINGLE)
        par = C.c_int(0)
        _lib.call("chmy_stokes_iterate", C.byref(g.desc()), _arr(self.V), _arr(self.tau), C.byref(self.Pr.desc()),
                  _arr(self.V2), _arr(self.tau2), C.byref(self.Pr2.desc()), _arr(self.tau_old),
                  C.byref(self.eta.desc()) if self.eta is not None else None,
                  C.byref(self.eta_ml.desc()) if self.eta_ml is not None else None,
                  C.byref(self.rhog.desc()) if self.rhog is not None else None,
                  _arr(self.f_V) if self.f_V is not None else None,
                  C.byref(self.f_P.desc()) if self.f_P is not None else None,
                  C.byref(self.divV.desc()), _arr(self.r_V), counts, comp, conds, int(niter), base,
                  base | (ITER_RESIDUAL if residual_last else 0), self.eta_s, self.eta_ve_s, self.G, self.dt,
                  self.p["dtau_Pr"], self.p["dtau_r"], self.p["nudtau"], C.byref(par), current_stream())
        if par.value == 1:
            self._swap()
        self.iters += niter

    def _pass2_call(self):
        """chmy_stokes_iter2: two PT iterations in one pass from (V, tau, Pr)
        into (V2, tau2, Pr2), the free-slip batch in-kernel between them."""
        g = self.grid
        counts, comp, conds, _keep = self._vbatch()
        _lib.call("chmy_stokes_iter2", C.byref(g.desc()), _arr(self.V), _arr(self.tau), C.byref(self.Pr.desc()),
                  _arr(self.V2), _arr(self.tau2), C.byref(self.Pr2.desc()), _arr(self.tau_old),
                  C.byref(self.eta.desc()) if self.eta is not None else None,
                  C.byref(self.eta_ml.desc()) if self.eta_ml is not None else None,
                  C.byref(self.rhog.desc()) if self.rhog is not None else None,
                  _arr(self.f_V) if self.f_V is not None else None,
                  C.byref(self.f_P.desc()) if self.f_P is not None else None,
                  counts, comp, conds, ITER_EXACT if self.exact else 0, self.eta_s, self.eta_ve_s, self.G,
                  self.dt, self.p["dtau_Pr"], self.p["dtau_r"], self.p["nudtau"], current_stream())

    def pass2(self):
        """Two PT iterations: one two-iteration pass, then the V batch on its
        output, then swap (the same state as two iteration() calls)."""
        self._pass2_call()
        self.apply_bc(self.V2)
        self._swap()
        self.iters += 2

    def pass2_first(self, V_first):
        """pass2() that also leaves the first iteration's V, after its V batch,
        in the VectorField V_first (chmy_stokes_iter2_first; FP64, no
        forcing): bitwise V after the first of two iteration() calls at every
        point either writes.  V_first's other points keep their values."""
        g = self.grid
        counts, comp, conds, _keep = self._vbatch()
        _lib.call("chmy_stokes_iter2_first", C.byref(g.desc()), _arr(self.V), _arr(self.tau),
                  C.byref(self.Pr.desc()), _arr(self.V2), _arr(self.tau2), C.byref(self.Pr2.desc()), _arr(V_first),
                  _arr(self.tau_old), C.byref(self.eta.desc()) if self.eta is not None else None,
                  C.byref(self.eta_ml.desc()) if self.eta_ml is not None else None,
                  C.byref(self.rhog.desc()) if self.rhog is not None else None,
                  counts, comp, conds, ITER_EXACT if self.exact else 0, self.eta_s, self.eta_ve_s, self.G,
                  self.dt, self.p["dtau_Pr"], self.p["dtau_r"], self.p["nudtau"], current_stream())
        self.apply_bc(V_first)
        self.apply_bc(self.V2)
        self._swap()
        self.iters += 2

    def iteration_unfused(self):
        """The reference's sequence: update_stress!, update_velocity! (per-point
        kernels, single buffers, r_V and ∇V written), then the V batch."""
        self.iteration_unfused_kernels()
        self.apply_bc(self.V)
        self.iters += 1

    def iteration_unfused_kernels(self):
        from .kernels import _box, update_stress, update_velocity, update_velocity_visc
        g = self.grid
        box = _box((0,) * self.nd, tuple(n + 2 for n in g.dims))
        p = self.p
        eta = self.eta if self.eta is not None else self.eta_s
        update_stress._fn(g, box, self.tau, self.Pr, self.divV, self.V, self.tau_old, eta, self.eta_ve_s, self.G,
                          self.dt, p["dtau_Pr"], p["dtau_r"], g)
        if self.eta_ml is not None:
            update_velocity_visc._fn(g, box, self.V, self.r_V, self.Pr, self.tau, self.rhog, self.eta_ml, self.G,
                                     self.dt, p["nudtau"], g)
        else:
            update_velocity._fn(g, box, self.V, self.r_V, self.Pr, self.tau, self.rhog, self.eta_ve_s,
                                p["nudtau"], g)

    def residual(self, tau_sc=1.0, lz=1.0, psc=1.0):
        """err = (max|∇V|·τsc, max|r_V.x|·lz/psc, ...) after zeroing boundary
        residual vertices (stokes_3d_inc_ve_T.jl:171-175).  Requires the last
        iteration to have written the residual fields."""
        names = self.grid.axes_names()
        key = id(self.r_V[0])
        rb = self._rbc.get(key)
        if rb is None:
            rb = batch(self.grid, *[(self.r_V[D], {names[D]: Dirichlet()}) for D in range(self.nd)])
            self._rbc[key] = rb
        rb.apply()
        from .kernels import maxabs_many
        m = maxabs_many([self.divV] + [self.r_V[D] for D in range(self.nd)])
        return [m[0] * tau_sc] + [m[1 + D] * lz / psc for D in range(self.nd)]

    def true_residual(self):
        """Residual norms of the current f64 state evaluated from scratch:
        [max|∇·V|, max|R_V| per component, max|R_τ|] (chmy_stokes_residual)."""
        if self.dtype != torch.float64:
            raise ValueError("true_residual needs an f64 state")
        if getattr(self, "_scr", None) is None:
            g = self.grid
            self._scr = (TensorField(self.arch, g, torch.float64), VectorField(self.arch, g, torch.float64),
                         Field(self.arch, g, Center(), torch.float64))
        toc, fV, fP = self._scr
        g = self.grid
        _lib.call("chmy_stokes_residual", C.byref(g.desc()), _arr(self.V), _arr(self.tau), C.byref(self.Pr.desc()),
                  _arr(self.tau_old), C.byref(self.eta.desc()) if self.eta is not None else None,
                  C.byref(self.rhog.desc()) if self.rhog is not None else None,
                  _arr(self.f_V) if self.f_V is not None else None, self.eta_s, self.G, self.dt,
                  _arr(toc), _arr(fV), C.byref(fP.desc()), C.byref(self.divV.desc()), _arr(self.r_V),
                  current_stream())
        from .kernels import maxabs
        err_tau = max(maxabs(t) for t in toc) / (self.G * self.dt)
        return self.residual() + [err_tau]

    def solve(self, eps=1e-8, max_iter=100000, ncheck=None, callback=None):
        """Iterate until all scaled residuals < eps (stokes_3d_inc_ve_T.jl:163-181).
        Raises on non-finite residuals."""
        ncheck = ncheck or 2 * self.grid.dims[0]
        err = None
        it = 0
        while it < max_iter:
            n = min(ncheck, max_iter - it)
            self.iterate(n, residual_last=True)
            it += n
            err = self.residual()
            if callback:
                callback(it, err)
            if not all(math.isfinite(e) for e in err):
                raise FloatingPointError(f"simulation failed, err = {err}")
            if all(e < eps for e in err):
                return it, err
        return max_iter, err


def inclusion_setup_3d(grid, *, eta0=1.0, eta_incl=1e-3, radius=0.1, T_amp=0.1, T_width=0.05):
    """Config: 3D VE Stokes+T with a spherical weak inclusion.  η on the full
    padded storage (logical -1..n+2) from the fma cell-centre coordinates:
    η = (|x| < r ? η_incl : η0)·exp(-T), T = T_amp·exp(-|x|²/T_width²)."""
    xs = [grid.coords(Center(), d, first=-1, count=grid.dims[d] + 4) for d in range(3)]
    X = xs[0][None, None, :]
    Y = xs[1][None, :, None]
    Z = xs[2][:, None, None]
    r2 = X ** 2 + Y ** 2 + Z ** 2
    T = T_amp * np.exp(-r2 / T_width ** 2)
    eta = np.where(r2 < radius ** 2, eta_incl, eta0) * np.exp(-T)
    return eta


def pure_shear_values(nd):
    """Normal boundary velocities ±1: x compressed (Vx = +1 | -1), last dim extended."""
    vals = {D: (0.0, 0.0) for D in range(nd)}
    vals[0] = (1.0, -1.0)
    vals[nd - 1] = (-1.0, 1.0)
    return vals


def pure_shear_initial(grid, V):
    """V_x = -2x, V_last = +2·x_last on every stored point (linear background)."""
    nd = grid.nd
    for D in (0, nd - 1):
        f = V[D]
        xs = grid.coords(f.loc, D, first=-1, count=f.dims[D] + 4)
        rate = -2.0 if D == 0 else 2.0
        shape = [1] * nd
        shape[nd - 1 - D] = -1
        f.data.copy_(torch.as_tensor(np.broadcast_to((rate * xs).reshape(shape), f.data.shape).copy()))


class MixedPTStokes:
    """FP32 PT sweeps with FP64 defect correction (build extension).

    forcing_dtype: storage of the sweeps' forcing (torch.bfloat16 by default,
    or torch.float32).

    The f64 master state x = (V, τ, Pr) lives in `self.x` (a PTStokes in f64;
    only its storage and BC batch are used).  Every `inner` iterations the f64
    residual R(x) is evaluated (chmy_stokes_residual) and becomes the forcing
    of an f32 PT solve for a correction d (`self.d`, homogeneous BCs), which
    starts from 0 and runs `inner` fused f32 iterations; then x += d
    (chmy_stokes_accumulate) and the real V batch is applied to x.  The
    convergence test uses the f64 residual of x: max|∇·V| and max|R_V| with
    boundary residual vertices zeroed (stokes_3d_inc_ve_T.jl:171-175).
    """

    def __init__(self, arch, grid, *, eta_field=None, eta=None, G=1.0, dt=1.0, normal_values=None,
                 re_m=2.3 * math.pi, r=0.5, rhog=None, inner=200, exact=False, forcing_dtype=torch.bfloat16):
        self.grid = grid
        self.inner = int(inner)
        self.x = PTStokes(arch, grid, dtype=torch.float64, eta_field=eta_field, eta=eta, G=G, dt=dt,
                          normal_values=normal_values, re_m=re_m, r=r, rhog=rhog, exact=exact)
        self.d = PTStokes(arch, grid, dtype=torch.float32, eta_field=eta_field, eta=eta, G=G, dt=dt,
                          normal_values={D: (0.0, 0.0) for D in range(grid.nd)}, re_m=re_m, r=r, exact=exact)
        # the correction's forcing (G dt R_tau in tau_old, R_V, R_P) is read once
        # per sweep and only sets the direction of each correction: bf16
        # storage halves its bytes; the f64 residual still sets the accuracy
        self.d.tau_old = TensorField(arch, grid, forcing_dtype)
        self.d.f_V = VectorField(arch, grid, forcing_dtype)
        self.d.f_P = Field(arch, grid, Center(), forcing_dtype)
        self.cycles = 0
        self.sweeps = 0

    def refresh(self):
        """f64 residual of x -> f32 forcing; returns the residual norms
        [max|∇·V|, max|R_V| per component, max|R_τ|] (τsc = lz = psc = 1).
        The stress residual is included because a state with τ = 0, Pr = 0 and
        a divergence-free linear V has ∇·V = 0 and R_V = 0 exactly."""
        x, d, g = self.x, self.d, self.grid
        _lib.call("chmy_stokes_residual", C.byref(g.desc()), _arr(x.V), _arr(x.tau), C.byref(x.Pr.desc()),
                  _arr(x.tau_old), C.byref(x.eta.desc()) if x.eta is not None else None,
                  C.byref(x.rhog.desc()) if x.rhog is not None else None,
                  _arr(x.f_V) if x.f_V is not None else None, x.eta_s, x.G, x.dt,
                  _arr(d.tau_old), _arr(d.f_V), C.byref(d.f_P.desc()), C.byref(x.divV.desc()), _arr(x.r_V),
                  current_stream())
        from .kernels import maxabs_many
        names = x.grid.axes_names()
        key = id(x.r_V[0])
        rb = x._rbc.get(key)
        if rb is None:
            rb = batch(x.grid, *[(x.r_V[D], {names[D]: Dirichlet()}) for D in range(x.nd)])
            x._rbc[key] = rb
        rb.apply()
        m = maxabs_many([x.divV] + [x.r_V[D] for D in range(x.nd)] + list(d.tau_old))
        err_tau = max(m[1 + x.nd:]) / (x.G * x.dt)
        return m[:1 + x.nd] + [err_tau]

    def set_eta(self, eta_field):
        self.x.set_eta(eta_field)
        self.d.set_eta(eta_field)

    def reset_state(self, V_init=None):
        self.x.reset_state(V_init)
        self.reset_correction()
        self.cycles = 0
        self.sweeps = 0

    def reset_correction(self):
        """d = 0: the set the next sweep reads.  The other set needs no reset:
        the sweep overwrites every point of it inside the Launcher box, and
        outside the box both sets hold the zeros they were allocated with,
        since no kernel and no homogeneous batch rule writes a non-zero there."""
        d = self.d
        for f in list(d.V) + list(d.tau) + [d.Pr]:
            f.data.zero_()

    def accumulate(self):
        x, d, g = self.x, self.d, self.grid
        _lib.call("chmy_stokes_accumulate", C.byref(g.desc()), _arr(x.V), _arr(x.tau), C.byref(x.Pr.desc()),
                  _arr(d.V), _arr(d.tau), C.byref(d.Pr.desc()), current_stream())
        x.apply_bc(x.V)

    def cycle(self):
        self.reset_correction()
        self.d.iterate(self.inner)
        self.accumulate()
        self.cycles += 1
        self.sweeps += self.inner

    def solve(self, eps=1e-8, max_cycles=10000, callback=None):
        """Defect-correction cycles until the f64 residual of x is below eps."""
        err = self.refresh()
        for c in range(max_cycles):
            if callback:
                callback(self.sweeps, err)
            if not all(math.isfinite(e) for e in err):
                raise FloatingPointError(f"simulation failed, err = {err}")
            if all(e < eps for e in err):
                return self.sweeps, err
            self.cycle()
            err = self.refresh()
        return self.sweeps, err

    def true_residual(self):
        return self.refresh()
