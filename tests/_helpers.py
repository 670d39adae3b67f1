"""Shared helpers: build matching product (device) and oracle (host) fields."""
import numpy as np
import torch

import chmy
import oracle as O

NP2TORCH = {np.float64: torch.float64, np.float32: torch.float32}


def locs_codes(f):
    return tuple(l.code for l in f.loc)


def ogrid(grid):
    """The oracle's grid for a product grid; a FunctionAxis hands the oracle
    its vertex function, which the oracle evaluates and derives from itself."""
    axes = getattr(grid, "axes", None)
    funcs = [getattr(a, "f", None) for a in axes] if axes else None
    return O.Grid(grid.origin, grid.extent, grid.dims, vertex_funcs=funcs)


def pair(arch, grid, og, loc, T=np.float64, fill=None, rng=None, lo=-1.0, hi=1.0):
    """A device Field and an oracle HField with identical full storage."""
    f = chmy.Field(arch, grid, loc, NP2TORCH[T])
    h = O.HField(og, locs_codes(f), T)
    if fill is not None:
        h.arr[...] = fill
    elif rng is not None:
        h.arr[...] = rng.uniform(lo, hi, h.arr.shape).astype(T)
    chmy.set_storage(f, h.arr)
    return f, h


def dev(f):
    return f.data.detach().cpu().numpy()


def assert_same(f, h, what="", interior_only=False, rtol=0.0):
    a = dev(f)
    b = h.arr
    if interior_only:
        sl = tuple(slice(2 * hh, 2 * hh + n) for hh, n in zip(h.halo, h.n))[::-1]
        a, b = a[sl], b[sl]
    if rtol == 0.0:
        if not np.array_equal(a, b, equal_nan=True):
            diff = np.nanmax(np.abs(a.astype(np.float64) - b.astype(np.float64)))
            bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
            raise AssertionError(f"{what}: not bitwise equal; max|diff|={diff:.3e}; "
                                 f"{len(bad)} mismatches, first at storage index {tuple(bad[0])}")
    else:
        scale = max(np.max(np.abs(b)), 1e-300)
        err = np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))) / scale
        assert err <= rtol, f"{what}: relative error {err:.3e} > {rtol:.1e}"


def vloc(D, nd):
    return chmy.vector_location(D, nd)
