"""Launcher (src/KernelLaunch.jl): iteration space, launch, optional bc!, sync.

`launch(arch, grid, (kernel, args), bc=batchset)` runs `kernel` over the
logical box I_d in [0, n_d+1] (worksize n .+ 2 at Offset(-1),
KernelLaunch.jl:41,105-126), then applies the boundary batch set, then
synchronizes (KernelLaunch.jl:117).  With `outer_width`, the box is split into
the inner box plus 2N outer slabs and each slab's side of the batch is applied
right after it (KernelLaunch.jl:152-183); on one GPU the slabs run in order on
one stream, so results equal the unsplit launch.
"""
from __future__ import annotations

from .architectures import synchronize


class Offset:
    def __init__(self, *o):
        self.o = tuple(o)

    def __add__(self, other):
        if isinstance(other, Offset):
            if len(self.o) == 1:
                return Offset(*(self.o[0] + x for x in other.o))
            return Offset(*(a + b for a, b in zip(self.o, other.o)))
        raise TypeError


class Launcher:
    def __init__(self, arch, grid, outer_width=None, sync=True):
        self.arch = arch
        self.grid = grid
        self.nd = grid.nd
        self.worksize = tuple(n + 2 for n in grid.dims)
        self.outer_width = None if outer_width is None else tuple(outer_width)
        self.sync = sync
        if self.outer_width is not None:
            if len(self.outer_width) != self.nd:
                raise ValueError("outer_width needs one entry per dim")
            if any(2 * w > s for w, s in zip(self.outer_width, self.worksize)):
                raise ValueError("outer_width too large for the grid")

    # boxes (KernelLaunch.jl:59-87)
    def inner_worksize(self):
        return tuple(s - 2 * w for s, w in zip(self.worksize, self.outer_width))

    def inner_offset(self):
        return self.outer_width

    def outer_worksize(self, D):
        return tuple(self.worksize[I] if I < D else (self.outer_width[I] if I == D else
                                                     self.worksize[I] - 2 * self.outer_width[I])
                     for I in range(self.nd))

    def outer_offset(self, D, S):
        return tuple(0 if I < D else ((0 if S == 1 else self.worksize[I] - self.outer_width[I]) if I == D
                                      else self.outer_width[I]) for I in range(self.nd))

    def __call__(self, arch, grid, kernel_and_args, bc=None):
        kernel, args = kernel_and_args
        base = -1  # Offset(-1)
        if self.outer_width is None:
            kernel.launch(grid, tuple(base + 1 for _ in range(self.nd)), self.worksize, *args)
            if bc is not None:
                bc.apply()
        else:
            lo = tuple(base + 1 + o for o in self.inner_offset())
            kernel.launch(grid, lo, self.inner_worksize(), *args)
            for D in reversed(range(self.nd)):
                for S in (1, 2):
                    lo = tuple(base + 1 + o for o in self.outer_offset(D, S))
                    ws = self.outer_worksize(D)
                    if all(w > 0 for w in ws):
                        kernel.launch(grid, lo, ws, *args)
                    if bc is not None:
                        bc.apply_side(D, S)
        if self.sync:
            synchronize(arch.backend if arch is not None else None)
