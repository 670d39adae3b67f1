"""FieldMask (src/GridOperators/field_mask.jl:1-117): one weight field per
staggered location, `at(ω, loc)` selecting the component that masked
operators multiply a field's knots by (masked_operators.jl:11-19).

1D: c, v; 2D: cc, vv, vc, cv; 3D: ccc, vvv, vcc, cvc, ccv, vvc, vcv, cvv
(x first), each a Field of the grid's size at that location, zero-filled like
the reference's.  The masked operators themselves are the kernels of
chmy.kernels given a FieldMask after their source field(s).
"""
from __future__ import annotations

import itertools

import torch

from .fields import Field
from .grids import Center, Vertex, expand_loc

_C, _V = Center(), Vertex()


def _name(loc):
    return "".join("v" if l.code else "c" for l in loc)


class FieldMask:
    """FieldMask(arch, grid, type=eltype(grid)) (field_mask.jl:1D :10-14,
    2D :30-37, 3D :71-82)."""

    def __init__(self, arch, grid, dtype=torch.float64, halo=1):
        self.grid = grid
        self.nd = grid.nd
        self.components = {}
        for loc in itertools.product((_C, _V), repeat=grid.nd):
            self.components[tuple(l.code for l in loc)] = Field(arch, grid, loc, dtype, halo)

    def at(self, loc):
        """at(ω, loc) (field_mask.jl:16-17, 39-42, 84-91): the component at loc."""
        codes = tuple(l.code for l in expand_loc(self.nd, loc))
        return self.components[codes]

    def __getattr__(self, name):
        # ω.cc, ω.vcv, ... as in the reference's field names
        comps = self.__dict__.get("components")
        if comps is not None and len(name) == self.__dict__.get("nd", -1) and set(name) <= {"c", "v"}:
            return comps[tuple(1 if ch == "v" else 0 for ch in name)]
        raise AttributeError(name)

    def __iter__(self):
        return iter(self.components.values())

    def fill_(self, value):
        """Every component to value (e.g. 1.0: the unmasked operators)."""
        for f in self:
            f.data.fill_(float(value))
        return self
