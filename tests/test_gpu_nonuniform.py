"""Grids with FunctionAxis axes (src/Grids/function_axis.jl) on the GPU against
the oracle, bitwise over the whole padded storage: every operator, plain and
masked, the bc! batches (one launch and per dim), and set! at the axis
coordinates, 1D-3D, FP64 and FP32, with stretched axes mixed with uniform
ones.  The metric tables the product builds on the host and reads on the
device must give the bits of the oracle, which derives centres, spacings and
weights from the same vertex function by itself."""
import math

import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy import Vertex as Vv
import oracle as O
from _helpers import NP2TORCH, assert_same, ogrid, pair
from test_gpu_kernels import _bc_specs, _ospec

pytestmark = pytest.mark.gpu
TYPES = [np.float64, np.float32]


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def stretched(o, L, n, a):
    return lambda i: o + L * ((i - 1) / n + a * math.sin(2 * math.pi * (i - 1) / n))


def grid(arch, nd):
    """x stretched; y uniform (2D/3D); z stretched (3D)."""
    if nd == 1:
        return chmy.StructuredGrid(arch, chmy.FunctionAxis(stretched(-1.0, 2.0, 37, 0.06), 37))
    if nd == 2:
        return chmy.StructuredGrid(arch, chmy.FunctionAxis(stretched(-1.0, 2.0, 70, 0.05), 70),
                                   chmy.UniformAxis(-0.5, 1.5, 33))
    return chmy.StructuredGrid(arch, chmy.FunctionAxis(stretched(-0.5, 1.0, 67, 0.04), 67),
                               chmy.UniformAxis(-0.5, 1.0, 13), chmy.FunctionAxis(stretched(-0.5, 1.2, 9, 0.07), 9))


def all_locs(nd):
    return [tuple(Cc() if (m >> d) & 1 == 0 else Vv() for d in range(nd)) for m in range(1 << nd)]


def flipped(loc, D):
    return tuple((Vv() if isinstance(l, Cc) else Cc()) if d == D else l for d, l in enumerate(loc))


def codes(loc):
    return tuple(l.code for l in loc)


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_operators_on_function_axes(arch, T, nd):
    g = grid(arch, nd)
    og = ogrid(g)
    rng = np.random.default_rng(70 + nd)
    L = chmy.Launcher(arch, g)
    K = chmy.kernels
    w = chmy.FieldMask(arch, g, dtype=NP2TORCH[T])
    hw = {}
    for c, f in w.components.items():
        h = O.HField(og, c, T)
        h.arr[...] = rng.random(h.arr.shape).astype(T)
        chmy.set_storage(f, h.arr)
        hw[c] = h
    for src in all_locs(nd):
        s, hs = pair(arch, g, og, src, T, rng=rng, lo=0.5, hi=2.0)
        k, hk = pair(arch, g, og, all_locs(nd)[-1], T, rng=rng, lo=0.5, hi=2.0)
        for D in range(nd):
            for name, kern in (("partial", K.partial), ("delta", K.delta), ("left", K.left), ("right", K.right)):
                d_, hd = pair(arch, g, og, flipped(src, D), T, fill=0)
                L(arch, g, (kern, (d_, s, D, g)))
                O.op(name, og, hd, hs, dim=D)
                assert_same(d_, hd, f"{name} dim {D} src {codes(src)}")
                L(arch, g, (kern, (d_, s, w, D, g)))
                O.op_masked(name, og, hd, hs, hw[codes(src)], dim=D)
                assert_same(d_, hd, f"masked {name} dim {D} src {codes(src)}")
            d_, hd = pair(arch, g, og, src, T, fill=0)
            L(arch, g, (K.partial2, (d_, s, D, g)))
            O.op("d2", og, hd, hs, dim=D)
            assert_same(d_, hd, f"∂² dim {D} {codes(src)}")
            L(arch, g, (K.partial_k_partial, (d_, s, k, w, D, g)))
            O.op_masked("dkd", og, hd, hs, hw[codes(src)], k=hk, dim=D)
            assert_same(d_, hd, f"masked ∂k∂ dim {D} {codes(src)}")
        d_, hd = pair(arch, g, og, src, T, fill=0)
        L(arch, g, (K.lapl, (d_, s, g)))
        O.op("lapl", og, hd, hs)
        assert_same(d_, hd, f"lapl {codes(src)}")
        L(arch, g, (K.divg_grad, (d_, s, k, g)))
        O.op("divg_grad", og, hd, hs, k=hk)
        assert_same(d_, hd, f"divg_grad {codes(src)}")
        for to in all_locs(nd):
            d_, hd = pair(arch, g, og, to, T, fill=0)
            L(arch, g, (K.lerp, (d_, s, g)))
            O.op("lerp", og, hd, hs)
            assert_same(d_, hd, f"lerp {codes(src)} -> {codes(to)}")
            L(arch, g, (K.hlerp, (d_, s, g)))
            O.op("hlerp", og, hd, hs)
            assert_same(d_, hd, f"hlerp {codes(src)} -> {codes(to)}")
    if nd >= 2:
        Vs = [pair(arch, g, og, chmy.vector_location(D, nd), T, rng=rng) for D in range(nd)]
        d_, hd = pair(arch, g, og, (Cc(),) * nd, T, fill=0)
        L(arch, g, (K.divg, (d_, [v for v, _ in Vs], g)))
        O.op("divg", og, hd, [h for _, h in Vs])
        assert_same(d_, hd, "divg")
        L(arch, g, (K.vmag, (d_, [v for v, _ in Vs], g)))
        O.op("vmag", og, hd, [h for _, h in Vs])
        assert_same(d_, hd, "vmag")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("nfields", [4, 9])
def test_bc_batches_on_function_axes(arch, T, nd, nfields):
    """Dirichlet (t = 2 Δc(I_f2) iΔv(I_f)) and Neumann (h = Δ(flip(loc), I_f))
    ghosts read from the axis tables; 4 fields take the one-launch batch set,
    9 the per-dim launches."""
    g = grid(arch, nd)
    og = ogrid(g)
    rng = np.random.default_rng(90 + nd)
    locs = all_locs(nd)
    for spec in _bc_specs(nd, None):
        fs = [pair(arch, g, og, locs[q % len(locs)], T, rng=rng) for q in range(nfields)]
        chmy.bc_(arch, g, *[(f, spec) for f, _ in fs])
        O.bc_apply(og, [(h, _ospec(spec, nd, {})) for _, h in fs])
        for q, (f, h) in enumerate(fs):
            assert_same(f, h, f"bc {spec} loc {codes(locs[q % len(locs)])}")


def test_set_at_function_axis_coordinates(arch):
    """set! and a FunctionField on a stretched grid evaluate at the axis's
    vertex and centre coordinates (function_axis.jl:9, abstract_axis.jl:18),
    bitwise against the host evaluation at the oracle's coordinates."""
    g = grid(arch, 3)
    og = ogrid(g)
    fn = lambda x, y, z: (x * y - z) / (1.5 + x * x)
    for loc in all_locs(3):
        f = chmy.Field(arch, g, loc)
        chmy.set_(f, fn, g)
        xs = np.meshgrid(*[og.coords(codes(loc), d) for d in range(3)], indexing="ij")
        assert np.array_equal(f.interior().cpu().numpy(), fn(*xs)), codes(loc)
    ff = chmy.FunctionField(fn, g, (Vv(), Cc(), Vv()))
    m = ff.materialize(arch)
    lo, ln = (-1, -1, -1), tuple(n + 4 for n in m.dims)
    xs = np.meshgrid(*[og.coords(codes(m.loc), d, first=lo[d], count=ln[d]) for d in range(3)], indexing="ij")
    assert np.array_equal(m.xyz().cpu().numpy(), fn(*xs))


def test_solver_kernels_refuse_function_axes(arch):
    """The example kernels and the solvers need uniform axes and say so (the
    reference's PT parameters read spacing(grid::UniformGrid) as well)."""
    from chmy.stokes import PTStokes
    g = grid(arch, 2)
    L = chmy.Launcher(arch, g)
    C = chmy.Field(arch, g, Cc())
    q = chmy.VectorField(arch, g)
    with pytest.raises(chmy.ChmyError, match="FunctionAxis"):
        L(arch, g, (chmy.kernels.compute_q, (q, C, 1.0, g)))
    with pytest.raises(ValueError, match="uniform"):
        PTStokes(arch, g, eta=1.0, G=1.0, dt=1.0)
