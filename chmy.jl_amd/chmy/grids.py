"""Locations, UniformGrid and StructuredGrid with FunctionAxis axes
(src/Grids/Grids.jl, uniform_axis.jl, function_axis.jl, abstract_axis.jl,
structured_grid.jl)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

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


class UniformAxis:
    """UniformAxis(origin, extent, length) (uniform_axis.jl:1-11)."""

    def __init__(self, origin, extent, length):
        self.origin, self.extent, self.length = float(origin), float(extent), int(length)


class FunctionAxis:
    """FunctionAxis(f, length) (function_axis.jl:1-9): vertex(i) = f(i) for
    every integer i, halo included; centres, spacings and weights follow from
    the vertices (abstract_axis.jl:18,26-29, interpolation.jl:29)."""

    def __init__(self, f, length):
        self.f, self.length = f, int(length)

    def vertex(self, i):
        return float(self.f(i))

    def center(self, i):
        return 0.5 * (self.vertex(i) + self.vertex(i + 1))

    def table(self):
        """The metric table of chmy_grid_t.axis: rows vertex, center, Δ(Center),
        Δ(Vertex), their inverses and the Center -> Vertex lerp weight, for
        logical i = -AXIS_PAD .. n+1+AXIS_PAD, in the reference's IEEE
        operations (center = 0.5 (v_i + v_i+1), inv = 1/x, weight
        (0.5 Δc(i-1)) iΔv(i))."""
        P, n = _lib.AXIS_PAD, self.length
        xv = np.array([self.vertex(k) for k in range(-P - 1, n + 3 + P)], dtype=np.float64)
        xc = 0.5 * (xv[:-1] + xv[1:])          # centres from -P-1
        dc = xv[1:] - xv[:-1]                  # Δ(Center, i) from -P-1
        dv = xc[1:] - xc[:-1]                  # Δ(Vertex, i) from -P
        L = n + 2 + 2 * P
        t = np.empty((_lib.AXIS_ROWS, L), dtype=np.float64)
        t[_lib.AX_VERTEX] = xv[1:L + 1]
        t[_lib.AX_CENTER] = xc[1:L + 1]
        t[_lib.AX_DC] = dc[1:L + 1]
        t[_lib.AX_DV] = dv[0:L]
        t[_lib.AX_IDC] = 1.0 / t[_lib.AX_DC]
        t[_lib.AX_IDV] = 1.0 / t[_lib.AX_DV]
        t[_lib.AX_WCV] = (0.5 * dc[0:L]) * t[_lib.AX_IDV]
        return t


class StructuredGrid(UniformGrid):
    """StructuredGrid(arch, axes...) (structured_grid.jl:5-11) with UniformAxis
    and FunctionAxis axes.  Each FunctionAxis's metric table lives on the
    device for as long as the grid; the operators, bc! and set! accept such a
    grid, the solver kernels need uniform axes."""

    def __init__(self, arch, *axes):
        origin = [a.origin if isinstance(a, UniformAxis) else a.vertex(1) for a in axes]
        extent = [a.extent if isinstance(a, UniformAxis) else a.vertex(a.length + 1) - a.vertex(1) for a in axes]
        super().__init__(arch, origin=origin, extent=extent, dims=[a.length for a in axes])
        self.axes = tuple(axes)
        device = arch.device if arch is not None else None
        self._tables = [None if isinstance(a, UniformAxis) else
                        torch.from_numpy(a.table()).to(device or "cpu").contiguous() for a in axes]
        self._spacing = tuple(s if isinstance(a, UniformAxis) else float("nan") for s, a in zip(self._spacing, axes))
        self._inv = tuple(s if isinstance(a, UniformAxis) else float("nan") for s, a in zip(self._inv, axes))

    def uniform(self):
        return all(t is None for t in self._tables)

    def desc(self):
        if self._desc is None:
            g = super().desc()
            for d, t in enumerate(self._tables):
                if t is not None:
                    g.axis[d] = t.data_ptr()
            self._desc = g
        return self._desc

    def coords(self, loc, dim, first=1, count=None):
        locs = expand_loc(self.nd, loc)
        if count is None:
            count = self.dims[dim] + locs[dim].code
        ax = self.axes[dim]
        if isinstance(ax, UniformAxis):
            g1 = UniformGrid(origin=(ax.origin,), extent=(ax.extent,), dims=(ax.length,))
            return g1.coords((locs[dim],), 0, first=first, count=count)
        fn = ax.vertex if locs[dim].code == _lib.VERTEX else ax.center
        return np.array([fn(i) for i in range(first, first + count)], dtype=np.float64)


def spacing(grid):
    """spacing(grid::UniformGrid) (structured_grid.jl:161)."""
    if any(s != s for s in grid._spacing):
        raise ValueError("spacing(grid) is defined for uniform axes only")
    return grid._spacing


def inv_spacing(grid):
    if any(s != s for s in grid._inv):
        raise ValueError("inv_spacing(grid) is defined for uniform axes only")
    return grid._inv


def coord(grid, loc, dim, i):
    return float(grid.coords(loc, dim, first=i, count=1)[0])
