"""Locations and UniformGrid (src/Grids/Grids.jl, uniform_axis.jl, structured_grid.jl)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class Location:
    code = -1

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))


class Center(Location):
    """Cell centre (src/Grids/Grids.jl)."""
    code = _lib.CENTER

    def __repr__(self):
        return "Center()"


class Vertex(Location):
    """Cell vertex / face (src/Grids/Grids.jl)."""
    code = _lib.VERTEX

    def __repr__(self):
        return "Vertex()"


def flip(loc):
    return Vertex() if isinstance(loc, Center) else Center()


def expand_loc(nd, loc):
    if isinstance(loc, Location):
        return (loc,) * nd
    loc = tuple(loc)
    if len(loc) != nd:
        raise ValueError(f"location tuple of length {len(loc)} for a {nd}-D grid")
    return loc


class UniformGrid:
    """UniformGrid(arch; origin, extent, dims) (structured_grid.jl:27-39).

    Per axis: spacing = extent/n and inv_spacing = inv(spacing)
    (uniform_axis.jl:7-12), computed in IEEE f64 as Julia does.
    """

    AXES = ("x", "y", "z")

    def __init__(self, arch=None, *, origin, extent, dims):
        if not (len(origin) == len(extent) == len(dims)) or not 1 <= len(dims) <= 3:
            raise ValueError("origin, extent and dims must have equal length 1..3")
        self.arch = arch
        self.nd = len(dims)
        self.origin = tuple(float(o) for o in origin)
        self.extent = tuple(float(e) for e in extent)
        self.dims = tuple(int(n) for n in dims)
        if any(n < 1 for n in self.dims):
            raise ValueError("dims must be >= 1")
        self._spacing = tuple(e / n for e, n in zip(self.extent, self.dims))
        self._inv = tuple(1.0 / s for s in self._spacing)
        self._desc = None

    def ndims(self):
        return self.nd

    def size(self, loc):
        """size(grid, loc): n cells for Center, n+1 for Vertex per dim."""
        locs = expand_loc(self.nd, loc)
        return tuple(n + l.code for n, l in zip(self.dims, locs))

    def axes_names(self):
        return self.AXES[: self.nd]

    def desc(self):
        if self._desc is None:
            g = _lib.ChmyGrid()
            g.nd = self.nd
            for d in range(self.nd):
                g.n[d] = self.dims[d]
                g.origin[d] = self.origin[d]
                g.extent[d] = self.extent[d]
                g.spacing[d] = self._spacing[d]
                g.inv_spacing[d] = self._inv[d]
            self._desc = g
        return self._desc

    def coords(self, loc, dim, first=1, count=None):
        """[coord(grid, loc, dim, i) for i in first:first+count-1]: vertex(i) =
        fma(i-1, Δ, o), center(i) = fma(i-1, Δ, fma(0.5, Δ, o))
        (uniform_axis.jl:18-19), evaluated by the library's host code."""
        locs = expand_loc(self.nd, loc)
        if count is None:
            count = self.dims[dim] + locs[dim].code
        out = np.empty(count, dtype=np.float64)
        _lib.call("chmy_coords", C.byref(self.desc()), dim, locs[dim].code, first, count,
                  out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def launcher_box(self):
        """Launcher iteration space: I_d in [0, n_d+1] (KernelLaunch.jl:41,109)."""
        return (0,) * self.nd, tuple(n + 2 for n in self.dims)


def spacing(grid):
    return grid._spacing


def inv_spacing(grid):
    return grid._inv


def coord(grid, loc, dim, i):
    return float(grid.coords(loc, dim, first=i, count=1)[0])
