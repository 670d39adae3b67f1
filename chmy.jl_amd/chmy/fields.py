"""Field storage (src/Fields/field.jl): padded layout, interior, set!, vector
and tensor fields.

A Field of logical size n (= size(grid, loc)) and halo H owns a zero-filled
device array of size n .+ 4H per dim, x fastest (a torch tensor of shape
reversed(n .+ 4H)); logical index I lives at storage I .+ 2H (1-based).
"""
from __future__ import annotations

from collections import namedtuple

import numpy as np
import torch

from . import _lib
from .grids import Center, Vertex, expand_loc

_DTYPES = {torch.float64: _lib.F64, torch.float32: _lib.F32, torch.bfloat16: _lib.BF16}


class Field:
    """Field{T,N,L,H} (field.jl:6-9), constructor field.jl:56-62."""

    def __init__(self, arch, grid, loc, dtype=torch.float64, halo=1, device=None):
        if dtype not in _DTYPES:
            raise TypeError(f"unsupported dtype {dtype}")
        self.grid = grid
        self.nd = grid.nd
        self.loc = expand_loc(grid.nd, loc)
        self.halo = (halo,) * grid.nd if isinstance(halo, int) else tuple(halo)
        self.dims = grid.size(self.loc)
        self.pad = tuple(n + 4 * h for n, h in zip(self.dims, self.halo))
        if device is None:
            device = arch.device if arch is not None else torch.device("cpu")
        self.data = torch.zeros(self.pad[::-1], dtype=dtype, device=device)
        self._desc = None

    # ------------------------------------------------------------------ API
    @property
    def dtype(self):
        return self.data.dtype

    def location(self):
        return self.loc

    def size(self):
        return self.dims

    def xyz(self):
        """Storage with Julia's index order (x first)."""
        return self.data.permute(*reversed(range(self.nd)))

    def interior(self, with_halo=False):
        """interior(f; with_halo) (field.jl:33-37) as a view."""
        sl = []
        for d in range(self.nd):
            h = self.halo[d]
            e = h if with_halo else 0
            sl.append(slice(2 * h - e, 2 * h + self.dims[d] + e))
        return self.xyz()[tuple(sl)]

    def desc(self):
        """chmy_field_t of this field (device storage only)."""
        if self.data.device.type != "cuda":
            raise _lib.ChmyError("field storage is on the host; kernels need device fields")
        if self._desc is None or self._desc.data != self.data.data_ptr():
            f = _lib.ChmyField()
            f.data = self.data.data_ptr()
            f.nd = self.nd
            f.dtype = _DTYPES[self.data.dtype]
            for d in range(self.nd):
                f.n[d] = self.dims[d]
                f.halo[d] = self.halo[d]
                f.loc[d] = self.loc[d].code
            self._desc = f
        return self._desc

    def swap(self, other):
        """Exchange storage with a field of the same shape (double buffering)."""
        if self.data.shape != other.data.shape or self.loc != other.loc:
            raise ValueError("swap needs fields of identical layout")
        self.data, other.data = other.data, self.data
        self._desc, other._desc = None, None

    def __repr__(self):
        return f"Field(loc={self.loc}, dims={self.dims}, halo={self.halo}, dtype={self.dtype})"


def interior(f, with_halo=False):
    return f.interior(with_halo=with_halo)


def location(f):
    return f.loc


def halo(f):
    return f.halo


def set_(f, value, grid=None, *, parameters=(), discrete=False):
    """set!(f, val | array | other | (grid, fun)) (field.jl:87-142) on the interior.

    A function (continuous: fun(x, y, z, *parameters) at the grid's fma
    coordinates; discrete: fun(grid, loc, ix, iy, iz, *parameters) at logical
    indices) is traced into a point program and evaluated on the GPU, as
    the reference's _set_continuous!/_set_discrete! kernels do; a
    FunctionField likewise (set!(f, other::AbstractField)).  On a host field
    (no GPU) the function is evaluated with numpy.  A VectorField/TensorField
    sets every component (field.jl:144-145)."""
    from .function_field import FunctionField, _params, trace
    if isinstance(f, tuple):
        for c in f:
            set_(c, value, grid, parameters=parameters, discrete=discrete)
        return
    it = f.interior()
    loc = f.loc
    if isinstance(value, FunctionField):
        # interior(other) into interior(f), index by index (field.jl:109-114):
        # the function at its own location
        if tuple(value.dims) != tuple(f.dims):
            raise ValueError(f"set_: FunctionField of size {value.dims} into a field of size {f.dims}")
        if f.data.is_cuda:
            value.program().run(value.grid, f)
            return
        parameters, discrete = value.parameters, value.discrete
        value, grid, loc = value.func, value.grid, value.loc
    if callable(value):
        if grid is None:
            raise ValueError("set_ with a function needs the grid")
        if f.data.is_cuda:
            trace(value, grid, f.loc, discrete=discrete, parameters=parameters).run(grid, f)
            return
        params = _params(parameters)
        if discrete:
            idx = np.meshgrid(*[np.arange(1, n + 1) for n in f.dims], indexing="ij")
            vals = value(grid, loc, *idx, *params)
        else:
            xs = np.meshgrid(*[grid.coords(loc, d) for d in range(f.nd)], indexing="ij")
            vals = value(*xs, *params)
        vals = np.array(np.broadcast_to(np.asarray(vals, dtype=np.float64), f.dims))
        it.copy_(torch.from_numpy(vals).to(it.dtype))
    elif isinstance(value, Field):
        it.copy_(value.interior())
    elif isinstance(value, (np.ndarray, torch.Tensor)):
        t = torch.as_tensor(np.asarray(value) if isinstance(value, np.ndarray) else value)
        it.copy_(t.to(dtype=it.dtype))
    else:
        it.fill_(float(value))


def set_storage(f, arr):
    """Copy a full padded storage array (shape reversed(n .+ 4H), x fastest) to f."""
    t = torch.as_tensor(np.ascontiguousarray(arr))
    if tuple(t.shape) != tuple(f.data.shape):
        raise ValueError(f"storage shape {tuple(t.shape)} != {tuple(f.data.shape)}")
    f.data.copy_(t.to(f.data.dtype))


def vector_location(D, nd):
    """Component D at Vertex along D, Center elsewhere (field.jl:148)."""
    return tuple(Vertex() if d == D else Center() for d in range(nd))


def VectorField(arch, grid, dtype=torch.float64, halo=1, device=None):
    names = grid.axes_names()
    NT = namedtuple("VectorField", names)
    return NT(*[Field(arch, grid, vector_location(D, grid.nd), dtype, halo, device) for D in range(grid.nd)])


def tensor_locations(nd):
    """TensorField component locations (field.jl:182-206)."""
    C, V = Center(), Vertex()
    if nd == 2:
        return {"xx": (C, C), "yy": (C, C), "xy": (V, V)}
    if nd == 3:
        return {"xx": (C, C, C), "yy": (C, C, C), "zz": (C, C, C),
                "xy": (V, V, C), "xz": (V, C, V), "yz": (C, V, V)}
    raise ValueError("TensorField needs a 2-D or 3-D grid")


def TensorField(arch, grid, dtype=torch.float64, halo=1, device=None):
    locs = tensor_locations(grid.nd)
    NT = namedtuple("TensorField", list(locs))
    return NT(*[Field(arch, grid, l, dtype, halo, device) for l in locs.values()])
