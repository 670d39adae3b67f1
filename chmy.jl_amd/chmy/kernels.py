"""HIP counterparts of the reference's Launcher kernels.

Each object stands where the reference's `@kernel` function stands in
`launch(arch, grid, kernel => args)`: the Launcher hands it a logical box and
the same argument list the reference kernel takes, and it calls the C ABI.

  compute_q, update_C              examples/diffusion_2d.jl:8-19
  diffusion_step                   both in one launch, double-buffered C (extension)
  update_old, update_stress,
  update_velocity, update_thermal_flux,
  update_thermal                   examples/stokes_3d_inc_ve_T.jl:11-77 (2D: stokes_2d_inc_ve_T.jl:11-60)
  thermal_step                     the last two in one launch, double-buffered T (extension)
  partial, lerp, divg, vmag, left,
  right                            GridOperators as used in test/test_grid_operators.jl
  delta, hlerp, partial2,
  partial_k_partial, lapl,
  divg_grad                        δ, hlerp, ∂², ∂k∂, lapl, divg_grad (partial_derivatives.jl,
                                   field_operators.jl, interpolation.jl); every operator but
                                   lerp/hlerp/vmag also masked, given a FieldMask
                                   (masked_operators.jl, masked_field_operators.jl)
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .architectures import current_stream
from .field_mask import FieldMask
from .fields import Field
from .function_field import FunctionField


def _box(lo, ws):
    b = _lib.ChmyBox()
    for d in range(len(lo)):
        b.lo[d] = lo[d]
        b.len[d] = ws[d]
    return b


def _arr(fields):
    fields = list(fields)
    return (_lib.ChmyField * len(fields))(*[f.desc() for f in fields])


def _opt(f):
    """An optional field argument; a FunctionField (the reference passes ρg as
    one, examples/stokes_2d_inc_ve_T.jl:107,146) is evaluated into a field."""
    if isinstance(f, FunctionField):
        f = f.materialize()
    return C.byref(f.desc()) if f is not None else None


class HipKernel:
    def __init__(self, name, fn):
        self.name = name
        self._fn = fn

    def launch(self, grid, lo, worksize, *args):
        self._fn(grid, _box(lo, worksize), *args)

    def __repr__(self):
        return f"HipKernel({self.name})"


def _compute_q(grid, box, q, Cf, chi, g):
    _lib.call("chmy_diffusion_flux", C.byref(grid.desc()), C.byref(box), _arr(q), C.byref(Cf.desc()),
              float(chi), current_stream())


def _update_C(grid, box, Cf, q, dt, g):
    _lib.call("chmy_diffusion_update", C.byref(grid.desc()), C.byref(box), C.byref(Cf.desc()), _arr(q),
              float(dt), current_stream())


def _diffusion_step(grid, box, Cf, C_out, q, chi, dt, g):
    _lib.call("chmy_diffusion_step", C.byref(grid.desc()), C.byref(box), C.byref(Cf.desc()), C.byref(C_out.desc()),
              _arr(q), float(chi), float(dt), current_stream())


def _update_old(grid, box, T, tau, T_old, tau_old):
    _lib.call("chmy_stokes_update_old", C.byref(grid.desc()), C.byref(box), _opt(T), _arr(tau), _opt(T_old),
              _arr(tau_old), current_stream())


def _update_stress(grid, box, tau, Pr, dV, V, tau_old, eta, eta_ve, G, dt, dtau_Pr, dtau_r, g):
    """eta: a number (reference form, with the given eta_ve) or a Center
    Field on its full padded storage (variable-viscosity extension)."""
    ef = eta if isinstance(eta, Field) else None
    es = 0.0 if ef is not None else float(eta)
    _lib.call("chmy_stokes_stress", C.byref(grid.desc()), C.byref(box), _arr(tau), C.byref(Pr.desc()),
              C.byref(dV.desc()), _arr(V), _arr(tau_old), _opt(ef), es, float(eta_ve), float(G), float(dt),
              float(dtau_Pr), float(dtau_r), current_stream())


def _update_velocity(grid, box, V, r_V, Pr, tau, rhog, eta_ve, nudtau, g):
    _lib.call("chmy_stokes_velocity", C.byref(grid.desc()), C.byref(box), _arr(V), _arr(r_V),
              C.byref(Pr.desc()), _arr(tau), _opt(rhog), None, float(eta_ve), 1.0, 1.0, float(nudtau),
              current_stream())


def _update_velocity_visc(grid, box, V, r_V, Pr, tau, rhog, eta_ml, G, dt, nudtau, g):
    _lib.call("chmy_stokes_velocity", C.byref(grid.desc()), C.byref(box), _arr(V), _arr(r_V),
              C.byref(Pr.desc()), _arr(tau), _opt(rhog), C.byref(eta_ml.desc()), 0.0, float(G), float(dt),
              float(nudtau), current_stream())


def _update_thermal_flux(grid, box, qT, T, V, lam, g):
    _lib.call("chmy_thermal_flux", C.byref(grid.desc()), C.byref(box), _arr(qT), C.byref(T.desc()), _arr(V),
              float(lam), current_stream())


def _update_thermal(grid, box, T, T_old, qT, dt, g):
    _lib.call("chmy_thermal_update", C.byref(grid.desc()), C.byref(box), C.byref(T.desc()),
              C.byref(T_old.desc()), _arr(qT), float(dt), current_stream())


def _thermal_step(grid, box, T, T_out, T_old, qT, V, lam, dt, g):
    _lib.call("chmy_thermal_step", C.byref(grid.desc()), C.byref(box), C.byref(T.desc()), C.byref(T_out.desc()),
              C.byref(T_old.desc()), _arr(qT), _arr(V), float(lam), float(dt), current_stream())


def _maxloc(grid, box, dst, src):
    _lib.call("chmy_maxloc", C.byref(grid.desc()), C.byref(box), C.byref(dst.desc()), C.byref(src.desc()),
              current_stream())


_PLAIN = {_lib.OP_PARTIAL, _lib.OP_LERP, _lib.OP_DIVG, _lib.OP_VMAG, _lib.OP_LEFT, _lib.OP_RIGHT}


def _op(code, has_k=False):
    def fn(grid, box, dst, src, *rest):
        # rest = ([k,] [ω,] dim, grid) for the per-dim operators, ([k,] [ω,] grid) otherwise:
        # the argument order of ∂x(f, ω, g, I), ∂k∂(f, k, ω, g, dim, I), divg(V, ω, g, I)
        srcs = list(src) if isinstance(src, (tuple, list)) else [src]
        rest = list(rest)
        k = rest.pop(0) if has_k else None
        w = rest.pop(0) if rest and isinstance(rest[0], FieldMask) else None
        dim = int(rest[0]) if len(rest) == 2 else 0
        if w is None and k is None and code in _PLAIN:
            _lib.call("chmy_op", code, C.byref(grid.desc()), C.byref(box), C.byref(dst.desc()), len(srcs),
                      _arr(srcs), dim, current_stream())
            return
        masks = _arr([w.at(s.loc) for s in srcs]) if w is not None else None
        _lib.call("chmy_op_masked", code, C.byref(grid.desc()), C.byref(box), C.byref(dst.desc()), len(srcs),
                  _arr(srcs), masks, C.byref(k.desc()) if k is not None else None, dim, current_stream())
    return fn


compute_q = HipKernel("compute_q!", _compute_q)
update_C = HipKernel("update_C!", _update_C)
# compute_q! + update_C! in one launch into a second buffer (build extension)
diffusion_step = HipKernel("compute_q! + update_C!", _diffusion_step)
update_old = HipKernel("update_old!", _update_old)
update_stress = HipKernel("update_stress!", _update_stress)
update_velocity = HipKernel("update_velocity!", _update_velocity)
update_velocity_visc = HipKernel("update_velocity! (variable viscosity)", _update_velocity_visc)
update_thermal_flux = HipKernel("update_thermal_flux!", _update_thermal_flux)
update_thermal = HipKernel("update_thermal!", _update_thermal)
thermal_step = HipKernel("update_thermal_flux! + update_thermal!", _thermal_step)
maxloc = HipKernel("maxloc", _maxloc)
# generic operators: (dst, src[, k][, ω][, dim], grid); ω a FieldMask gives the
# masked operator (masked_operators.jl, masked_field_operators.jl)
partial = HipKernel("∂", _op(_lib.OP_PARTIAL))
delta = HipKernel("δ", _op(_lib.OP_DELTA))
lerp = HipKernel("lerp", _op(_lib.OP_LERP))
hlerp = HipKernel("hlerp", _op(_lib.OP_HLERP))
divg = HipKernel("divg", _op(_lib.OP_DIVG))
vmag = HipKernel("vmag", _op(_lib.OP_VMAG))
left = HipKernel("left", _op(_lib.OP_LEFT))
right = HipKernel("right", _op(_lib.OP_RIGHT))
partial2 = HipKernel("∂²", _op(_lib.OP_D2))
partial_k_partial = HipKernel("∂k∂", _op(_lib.OP_DKD, has_k=True))
lapl = HipKernel("lapl", _op(_lib.OP_LAPL))
divg_grad = HipKernel("divg_grad", _op(_lib.OP_DIVG_GRAD, has_k=True))


def maxabs(f, box=None):
    """maximum(abs.(interior(f))) on the device, f64 result on the host."""
    out = C.c_double()
    b = _box((1,) * f.nd, f.dims) if box is None else box
    _lib.call("chmy_maxabs", C.byref(f.desc()), C.byref(b), C.byref(out), current_stream())
    return out.value


_MAX_BUF = {}


def maxabs_many(fields):
    """maximum(abs.(interior(f))) of each field: the reductions run on the
    device into one buffer and the host reads them once (one synchronisation
    for a whole convergence check)."""
    import torch
    n = len(fields)
    if n == 0:
        return []
    dev = fields[0].data.device
    buf = _MAX_BUF.get(dev)
    if buf is None or buf.numel() < n:
        buf = torch.empty(max(n, 16), dtype=torch.float64, device=dev)
        _MAX_BUF[dev] = buf
    for q, f in enumerate(fields):
        b = _box((1,) * f.nd, f.dims)
        _lib.call("chmy_maxabs_dev", C.byref(f.desc()), C.byref(b),
                  C.cast(C.c_void_p(buf.data_ptr() + 8 * q), C.POINTER(C.c_double)), current_stream())
    return buf[:n].cpu().tolist()


def sumsq(f, box=None):
    out = C.c_double()
    b = _box((1,) * f.nd, f.dims) if box is None else box
    _lib.call("chmy_sumsq", C.byref(f.desc()), C.byref(b), C.byref(out), current_stream())
    return out.value
