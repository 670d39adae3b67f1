"""Masked operators (FieldMask; src/GridOperators/masked_operators.jl,
masked_field_operators.jl, field_mask.jl) and the second-order and harmonic
operators (δ, ∂², ∂k∂, lapl, divg_grad, hlerp) on the GPU against the oracle,
bitwise over the whole padded storage, 1D-3D, FP64 and FP32, every source
location and dim, through the Launcher (box [0, n+1])."""
import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy import Vertex as Vv
import oracle as O
from _helpers import NP2TORCH, assert_same, ogrid, pair

pytestmark = pytest.mark.gpu
TYPES = [np.float64, np.float32]


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def grids(nd):
    if nd == 1:
        return chmy.UniformGrid(origin=(-1.0,), extent=(2.0,), dims=(37,))
    if nd == 2:
        return chmy.UniformGrid(origin=(-1.0, -0.5), extent=(2.0, 1.5), dims=(70, 33))
    return chmy.UniformGrid(origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.0, 1.2), dims=(67, 13, 9))


def all_locs(nd):
    return [tuple(Cc() if (m >> d) & 1 == 0 else Vv() for d in range(nd)) for m in range(1 << nd)]


def flipped(loc, D):
    return tuple((Vv() if isinstance(l, Cc) else Cc()) if d == D else l for d, l in enumerate(loc))


def mask_pair(arch, g, og, T, rng, kind):
    """A device FieldMask and its oracle components (by location codes)."""
    w = chmy.FieldMask(arch, g, dtype=NP2TORCH[T])
    host = {}
    for codes, f in w.components.items():
        h = O.HField(og, codes, T)
        if kind == "indicator":
            h.arr[...] = (rng.random(h.arr.shape) < 0.7).astype(T)
        else:
            h.arr[...] = rng.random(h.arr.shape).astype(T)
        chmy.set_storage(f, h.arr)
        host[codes] = h
    return w, host


def codes(loc):
    return tuple(l.code for l in loc)


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("kind", ["indicator", "fraction"])
def test_masked_per_dim_operators(arch, T, nd, kind):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(20 + nd)
    L = chmy.Launcher(arch, g)
    w, hw = mask_pair(arch, g, og, T, rng, kind)
    K = chmy.kernels
    for src_loc in all_locs(nd):
        s, hs = pair(arch, g, og, src_loc, T, rng=rng)
        k, hk = pair(arch, g, og, src_loc, T, rng=rng, lo=0.5, hi=2.0)
        hm = hw[codes(src_loc)]
        for D in range(nd):
            out = flipped(src_loc, D)
            for name, kern in (("partial", K.partial), ("delta", K.delta), ("left", K.left), ("right", K.right)):
                d_, hd = pair(arch, g, og, out, T, fill=0)
                L(arch, g, (kern, (d_, s, w, D, g)))
                O.op_masked(name, og, hd, hs, hm, dim=D)
                assert_same(d_, hd, f"masked {name} dim {D} src {codes(src_loc)}")
            d_, hd = pair(arch, g, og, src_loc, T, fill=0)
            L(arch, g, (K.partial2, (d_, s, w, D, g)))
            O.op_masked("d2", og, hd, hs, hm, dim=D)
            assert_same(d_, hd, f"masked ∂² dim {D} src {codes(src_loc)}")
            L(arch, g, (K.partial_k_partial, (d_, s, k, w, D, g)))
            O.op_masked("dkd", og, hd, hs, hm, k=hk, dim=D)
            assert_same(d_, hd, f"masked ∂k∂ dim {D} src {codes(src_loc)}")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_masked_divg_lapl_divg_grad(arch, T, nd):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(40 + nd)
    L = chmy.Launcher(arch, g)
    w, hw = mask_pair(arch, g, og, T, rng, "fraction")
    K = chmy.kernels
    Vs = [pair(arch, g, og, chmy.vector_location(D, nd), T, rng=rng) for D in range(nd)]
    d_, hd = pair(arch, g, og, (Cc(),) * nd, T, fill=0)
    L(arch, g, (K.divg, (d_, [v for v, _ in Vs], w, g)))
    O.op_masked("divg", og, hd, [h for _, h in Vs], [hw[codes(chmy.vector_location(D, nd))] for D in range(nd)])
    assert_same(d_, hd, "masked divg")
    for src_loc in all_locs(nd):
        s, hs = pair(arch, g, og, src_loc, T, rng=rng)
        k, hk = pair(arch, g, og, (Cc(),) * nd, T, rng=rng, lo=0.5, hi=2.0)   # k at centres, lerped
        d_, hd = pair(arch, g, og, src_loc, T, fill=0)
        L(arch, g, (K.lapl, (d_, s, w, g)))
        O.op_masked("lapl", og, hd, hs, hw[codes(src_loc)])
        assert_same(d_, hd, f"masked lapl {codes(src_loc)}")
        L(arch, g, (K.divg_grad, (d_, s, k, w, g)))
        O.op_masked("divg_grad", og, hd, hs, hw[codes(src_loc)], k=hk)
        assert_same(d_, hd, f"masked divg_grad {codes(src_loc)}")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_second_order_and_harmonic_operators(arch, T, nd):
    """Unmasked δ, ∂², ∂k∂, lapl, divg_grad and hlerp against the oracle."""
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(60 + nd)
    L = chmy.Launcher(arch, g)
    K = chmy.kernels
    for src_loc in all_locs(nd):
        s, hs = pair(arch, g, og, src_loc, T, rng=rng, lo=0.5, hi=2.0)     # positive: hlerp takes inverses
        k, hk = pair(arch, g, og, all_locs(nd)[-1], T, rng=rng, lo=0.5, hi=2.0)   # k at vertices
        for D in range(nd):
            d_, hd = pair(arch, g, og, flipped(src_loc, D), T, fill=0)
            L(arch, g, (K.delta, (d_, s, D, g)))
            O.op("delta", og, hd, hs, dim=D)
            assert_same(d_, hd, f"δ dim {D} {codes(src_loc)}")
            d_, hd = pair(arch, g, og, src_loc, T, fill=0)
            L(arch, g, (K.partial2, (d_, s, D, g)))
            O.op("d2", og, hd, hs, dim=D)
            assert_same(d_, hd, f"∂² dim {D} {codes(src_loc)}")
            L(arch, g, (K.partial_k_partial, (d_, s, k, D, g)))
            O.op("dkd", og, hd, hs, k=hk, dim=D)
            assert_same(d_, hd, f"∂k∂ dim {D} {codes(src_loc)}")
        d_, hd = pair(arch, g, og, src_loc, T, fill=0)
        L(arch, g, (K.lapl, (d_, s, g)))
        O.op("lapl", og, hd, hs)
        assert_same(d_, hd, f"lapl {codes(src_loc)}")
        L(arch, g, (K.divg_grad, (d_, s, k, g)))
        O.op("divg_grad", og, hd, hs, k=hk)
        assert_same(d_, hd, f"divg_grad {codes(src_loc)}")
        for to in all_locs(nd):
            d_, hd = pair(arch, g, og, to, T, fill=0)
            L(arch, g, (K.hlerp, (d_, s, g)))
            O.op("hlerp", og, hd, hs)
            assert_same(d_, hd, f"hlerp {codes(src_loc)} -> {codes(to)}")


def test_unit_mask_gives_the_plain_operators(arch):
    """ω ≡ 1: the masked kernels equal the plain ones bitwise (2D, FP64)."""
    g = grids(2)
    og = ogrid(g)
    rng = np.random.default_rng(5)
    L = chmy.Launcher(arch, g)
    w = chmy.FieldMask(arch, g).fill_(1.0)
    s, _ = pair(arch, g, og, (Cc(), Vv()), np.float64, rng=rng)
    a, _ = pair(arch, g, og, (Vv(), Vv()), np.float64, fill=0)
    b, _ = pair(arch, g, og, (Vv(), Vv()), np.float64, fill=0)
    L(arch, g, (chmy.kernels.partial, (a, s, w, 0, g)))
    L(arch, g, (chmy.kernels.partial, (b, s, 0, g)))
    assert bool((a.data == b.data).all())


def test_masked_argument_errors(arch):
    g = grids(2)
    og = ogrid(g)
    L = chmy.Launcher(arch, g)
    w = chmy.FieldMask(arch, g)
    s, _ = pair(arch, g, og, (Cc(), Cc()), np.float64, fill=1.0)
    d_, _ = pair(arch, g, og, (Cc(), Cc()), np.float64, fill=0)
    with pytest.raises(chmy.ChmyError, match="no masked form"):
        L(arch, g, (chmy.kernels.lerp, (d_, s, w, g)))
    w32 = chmy.FieldMask(arch, g, dtype=torch.float32)
    with pytest.raises(chmy.ChmyError, match="dtype"):
        L(arch, g, (chmy.kernels.partial, (d_, s, w32, 0, g)))
