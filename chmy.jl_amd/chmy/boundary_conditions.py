"""bc! batches (src/BoundaryConditions/batch.jl, first_order_boundary_condition.jl,
boundary_function.jl).

A per-field condition is a FirstOrderBC (all dims and sides) or a dict
{"x": bc | (bc_left, bc_right), ...} (the reference's NamedTuple form; missing
dims are `nothing`).  batch() normalises a list of (field, condition) pairs
into one FieldBatch per (dim, side), pruning fields without a condition
(batch.jl:72-157).  bc_() applies the set on the device: dims N..1, side 1 then
side 2, fields in batch order (batch.jl:20-29).

A condition's value is None (zero), a number, a Field over the tangential
indices, or a BoundaryFunction (a plain function is taken as a continuous one
over the reduced dimensions, boundary_function.jl:78-81).  batch() evaluates
a BoundaryFunction once per (field, dim, side) on the host, at the points
first_order_boundary_condition.jl:42-84 evaluates it, into a value Field over
the tangential range 0..n+2 that the device batch reads.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .architectures import current_stream
from .fields import Field, set_storage
from .grids import Center, UniformGrid, coord, flip


class FirstOrderBC:
    """FirstOrderBC{T,Kind} (first_order_boundary_condition.jl:9-32)."""

    def __init__(self, kind, value=None):
        self.kind = kind
        self.value = value

    def __repr__(self):
        name = "Dirichlet" if self.kind == _lib.DIRICHLET else "Neumann"
        return f"{name}({0 if self.value is None else self.value})"

    def desc(self, keep):
        b = _lib.ChmyBC()
        b.kind = self.kind
        v = self.value
        if v is None:
            b.vkind = _lib.VALUE_ZERO
        elif isinstance(v, Field):
            b.vkind = _lib.VALUE_FIELD
            fd = v.desc()
            keep.append(fd)
            b.vfield = C.pointer(fd)
        else:
            b.vkind = _lib.VALUE_SCALAR
            b.value = float(v)
        return b


class BoundaryFunction:
    """BoundaryFunction(fun; discrete, parameters, reduce_dims)
    (boundary_function.jl:46-75).  Continuous: fun(coordinates..., parameters...)
    of the evaluation point, without the boundary's own dim when reduce_dims.
    Discrete: fun(grid, loc, dim, indices..., parameters...) with Chmy's logical
    indices (interior from 1), likewise reduced."""

    def __init__(self, fun, discrete=False, parameters=None, reduce_dims=True):
        self.fun = fun
        self.discrete = bool(discrete)
        self.parameters = () if parameters is None else tuple(parameters)
        self.reduce_dims = bool(reduce_dims)

    def __call__(self, grid, loc, dim, I):
        """The value at logical index I (all dims) for location tuple loc
        (boundary_function.jl:34-44)."""
        keep = [d for d in range(grid.nd) if not (self.reduce_dims and d == dim)]
        if self.discrete:
            return self.fun(grid, loc, dim, *[I[d] for d in keep], *self.parameters)
        return self.fun(*[coord(grid, loc[d], d, I[d]) for d in keep], *self.parameters)

    def __repr__(self):
        return f"BoundaryFunction({getattr(self.fun, '__name__', self.fun)})"


def _as_value(value):
    if value is None or isinstance(value, (Field, BoundaryFunction)) or not callable(value):
        return value
    return BoundaryFunction(value)   # FirstOrderBC{Kind}(f::Function) (boundary_function.jl:78-81)


def Dirichlet(value=None):
    return FirstOrderBC(_lib.DIRICHLET, _as_value(value))


def Neumann(value=None):
    return FirstOrderBC(_lib.NEUMANN, _as_value(value))


def _materialise(grid, f, d, side, cond):
    """A BoundaryFunction condition of field f on dim d, side (0 or 1) as a
    condition with a value Field (or a number in 1D).  The point and location
    are those of first_order_boundary_condition.jl: a Vertex-located Dirichlet
    is evaluated at the boundary node (1 or m) with f's location (:50-54), a
    Center-located Dirichlet at the ghost (0 or m+1) with Center (:59-71), a
    Neumann at the ghost with the flipped location (:76-84); tangential indices
    run over 0..n+2."""
    bf = cond.value
    nd = grid.nd
    m = f.dims[d]
    vertex = f.loc[d].code == _lib.VERTEX
    if cond.kind == _lib.DIRICHLET and vertex:
        loc, normal = f.loc, (1 if side == 0 else m)
    elif cond.kind == _lib.DIRICHLET:
        loc, normal = (Center(),) * nd, (0 if side == 0 else m + 1)
    else:
        loc, normal = tuple(flip(l) for l in f.loc), (0 if side == 0 else m + 1)
    if nd == 1:
        return FirstOrderBC(cond.kind, float(bf(grid, loc, d, (normal,))))
    tang = [t for t in range(nd) if t != d]
    tg = UniformGrid(origin=[grid.origin[t] for t in tang], extent=[grid.extent[t] for t in tang],
                     dims=[grid.dims[t] for t in tang])
    vf = Field(None, tg, Center(), f.dtype, device=f.data.device)
    arr = np.zeros(tuple(reversed(vf.pad)))
    ranges = [range(0, grid.dims[t] + 3) for t in tang]
    for idx in np.ndindex(*[len(r) for r in ranges]):
        I = [0] * nd
        I[d] = normal
        for q, t in enumerate(tang):
            I[t] = ranges[q][idx[q]]
        # logical tangential index i sits at storage i + 1 of the value field
        arr[tuple(I[t] + 1 for t in reversed(tang))] = bf(grid, loc, d, I)
    set_storage(vf, arr)
    return FirstOrderBC(cond.kind, vf)


class EmptyBatch:
    fields = ()
    conditions = ()


class FieldBatch:
    """FieldBatch{K} (batch.jl:45-52): K fields with one condition each."""

    def __init__(self, fields, conditions):
        self.fields = tuple(fields)
        self.conditions = tuple(conditions)

    def __repr__(self):
        return f"FieldBatch ({len(self.fields)} fields)"


def _regularise(nd, names, spec):
    """regularise_impl (batch.jl:135-136): per dim, per side condition or None."""
    if spec is None:
        return [(None, None)] * nd
    if isinstance(spec, FirstOrderBC):
        return [(spec, spec)] * nd
    if isinstance(spec, dict):
        out = []
        for d in range(nd):
            b = spec.get(names[d], spec.get(d))
            if isinstance(b, tuple):
                out.append((b[0], b[1]))
            else:
                out.append((b, b))
        unknown = [k for k in spec if k not in names and k not in range(nd)]
        if unknown:
            raise ValueError(f"unknown dims {unknown} in boundary condition")
        return out
    raise TypeError(f"unsupported boundary condition {spec!r}")


class BatchSet:
    """NTuple{N, (side-1 batch, side-2 batch)} with cached C descriptors."""

    def __init__(self, grid, per_dim):
        self.grid = grid
        self.per_dim = per_dim
        self._c = None

    def __getitem__(self, d):
        return self.per_dim[d]

    def _key(self):
        """Storage pointers the cached descriptors were built from (fields swap
        storage under double buffering)."""
        key = []
        for d in range(self.grid.nd):
            for s in range(2):
                b = self.per_dim[d][s]
                for f, cond in zip(b.fields, b.conditions):
                    key.append(f.data.data_ptr())
                    if isinstance(cond.value, Field):
                        key.append(cond.value.data.data_ptr())
        return tuple(key)

    def _descriptors(self):
        key = self._key()
        if self._c is None:
            self._c = {}
        hit = self._c.get(key)
        if hit is None:
            if len(self._c) >= 4:          # double buffering alternates two keys
                self._c.clear()
            keep = []
            counts = (C.c_int32 * 6)()
            fl, bl = [], []
            for d in range(self.grid.nd):
                for s in range(2):
                    b = self.per_dim[d][s]
                    counts[2 * d + s] = len(b.fields)
                    for f, cond in zip(b.fields, b.conditions):
                        fl.append(f.desc())
                        bl.append(cond.desc(keep))
            n = max(1, len(fl))
            fa = (_lib.ChmyField * n)(*fl)
            ba = (_lib.ChmyBC * n)(*bl)
            hit = (counts, fa, ba, keep)
            self._c[key] = hit
        return hit

    def apply(self, stream=None):
        counts, fa, ba, _keep = self._descriptors()
        _lib.call("chmy_bc_batchset", C.byref(self.grid.desc()), counts, fa, ba,
                  current_stream() if stream is None else stream)

    def apply_side(self, d, s, stream=None):
        """bc!(Side(s), Dim(d), arch, grid, batch[d][s]) — one side of one dim."""
        b = self.per_dim[d][s - 1]
        if not b.fields:
            return
        keep = []
        fa = (_lib.ChmyField * len(b.fields))(*[f.desc() for f in b.fields])
        ba = (_lib.ChmyBC * len(b.fields))(*[c.desc(keep) for c in b.conditions])
        _lib.call("chmy_bc_side", C.byref(self.grid.desc()), d, s, len(b.fields), fa, ba,
                  current_stream() if stream is None else stream)


def batch(grid, *pairs, exchange=None):
    """batch(grid, f => bc, ...; exchange) (batch.jl:72-76).  `exchange` names
    fields to exchange with neighbouring ranks; on a single-device Bounded
    grid it is a no-op (batch.jl:103-105)."""
    names = grid.axes_names()
    regs = []
    for f, spec in pairs:
        r = _regularise(grid.nd, names, spec)
        r = [tuple(_materialise(grid, f, d, s, c) if c is not None and isinstance(c.value, BoundaryFunction) else c
                   for s, c in enumerate(r[d])) for d in range(grid.nd)]
        regs.append((f, r))
    per_dim = []
    for d in range(grid.nd):
        sides = []
        for s in range(2):
            fs = [f for f, r in regs if r[d][s] is not None]
            cs = [r[d][s] for f, r in regs if r[d][s] is not None]
            sides.append(FieldBatch(fs, cs) if fs else EmptyBatch())
        per_dim.append(tuple(sides))
    return BatchSet(grid, per_dim)


def bc_(arch, grid, *pairs_or_set, exchange=None):
    """bc!(arch, grid, f => bc, ...) or bc!(arch, grid, batchset) (batch.jl:20-29,157)."""
    if len(pairs_or_set) == 1 and isinstance(pairs_or_set[0], BatchSet):
        bs = pairs_or_set[0]
    else:
        bs = batch(grid, *pairs_or_set, exchange=exchange)
    bs.apply()
    return bs
