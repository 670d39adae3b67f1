"""Pins of the oracle's FunctionAxis support (src/Grids/function_axis.jl,
abstract_axis.jl:18,26-29; interpolation.jl:28-29; partial_derivatives.jl:5;
first_order_boundary_condition.jl:59-84).

- A FunctionAxis whose vertices are exactly uniform (dyadic spacing) has
  exactly the uniform centres, spacings and weights, so every operator and
  boundary condition must give the bits of the uniform grid, which the
  reference's own tests pin.
- On a strongly non-uniform axis the Center -> Vertex weight
  0.5 Δc(i-1) iΔv(i) makes lerp reproduce affine (and, as a tensor product,
  bilinear) fields at the vertex coordinates: a weight taken at the wrong
  index fails it.
- ∂ scales δ by iΔ at the source field's location and the output index, the
  Center Dirichlet ghost uses t = 2 Δc(I_f2) iΔv(I_f) and the Neumann ghost
  h = Δ(flip(loc), I_f): the reference's indices, which on a non-uniform axis
  differ from the geometric ones (the derivative of a Center field at vertex
  I divides by the width of cell I, not by the distance of the two centres).
  These tests compute the spacings from the vertex function themselves and
  fix which location and index each formula reads; the mistakes planted in
  the oracle (the other location, swapped indices) fail them.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

C_, V_ = O.CENTER, O.VERTEX


def fma(a, b, c):
    """a*b + c rounded once (Fraction arithmetic is exact, its float() rounds to nearest)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def stretched(n, a=0.05):
    """A monotone non-uniform vertex function on [0, 1] (and beyond, for the halo)."""
    return lambda i: (i - 1) / n + a * math.sin(2 * math.pi * (i - 1) / n)


def dyadic(o, h):
    return lambda i: o + (i - 1) * h


def _rand(g, loc, rng, lo=-1.0, hi=1.0):
    f = O.HField(g, loc)
    f.arr[...] = rng.uniform(lo, hi, f.arr.shape)
    return f


def _locs(nd):
    return [tuple((m >> d) & 1 for d in range(nd)) for m in range(1 << nd)]


def _flip(loc, D):
    return tuple(1 - l if d == D else l for d, l in enumerate(loc))


GRIDS = {1: ((8,), (-0.5,), (1.0,)), 2: ((8, 4), (-0.5, 0.25), (1.0, 0.5)), 3: ((8, 4, 4), (-0.5, 0.25, 0.0), (1.0, 0.5, 2.0))}


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_dyadic_function_axes_give_the_uniform_bits(nd):
    dims, o, e = GRIDS[nd]
    gu = O.Grid(o, e, dims)
    gf = O.Grid(o, e, dims, vertex_funcs=[dyadic(o[d], e[d] / dims[d]) for d in range(nd)])
    rng = np.random.default_rng(nd)
    for loc in _locs(nd):
        f = _rand(gu, loc, rng, 0.5, 2.0)
        k = _rand(gu, loc, rng, 0.5, 2.0)
        for d in range(nd):
            assert np.array_equal(gu.coords(loc, d, first=-1, count=dims[d] + 4),
                                  gf.coords(loc, d, first=-1, count=dims[d] + 4))
        for name in ("partial", "delta", "d2", "dkd"):
            for D in range(nd):
                to = loc if name in ("d2", "dkd") else _flip(loc, D)
                a, b = O.HField(gu, to), O.HField(gf, to)
                O.op(name, gu, a, f, k=k if name == "dkd" else None, dim=D)
                O.op(name, gf, b, f, k=k if name == "dkd" else None, dim=D)
                assert np.array_equal(a.arr, b.arr), (name, loc, D)
        for name in ("lerp", "hlerp"):
            for to in _locs(nd):
                a, b = O.HField(gu, to), O.HField(gf, to)
                O.op(name, gu, a, f)
                O.op(name, gf, b, f)
                assert np.array_equal(a.arr, b.arr), (name, loc, to)
        for name in ("lapl", "divg_grad"):
            a, b = O.HField(gu, loc), O.HField(gf, loc)
            O.op(name, gu, a, f, k=k if name == "divg_grad" else None)
            O.op(name, gf, b, f, k=k if name == "divg_grad" else None)
            assert np.array_equal(a.arr, b.arr), (name, loc)
        for bc in (O.Dirichlet(0.7), O.Neumann(-0.3)):
            a = O.HField(gu, loc)
            b = O.HField(gf, loc)
            a.arr[...] = f.arr
            b.arr[...] = f.arr
            O.bc_apply(gu, [(a, bc)])
            O.bc_apply(gf, [(b, bc)])
            assert np.array_equal(a.arr, b.arr), (bc, loc)


@pytest.mark.parametrize("nd", [1, 2])
def test_lerp_reproduces_affine_and_bilinear_fields(nd):
    """lerp to every location of a + b x (+ c y + d x y in 2D) sampled on a
    stretched axis equals the function at the target coordinates (to a few
    ulp): the weights are those of linear interpolation in x."""
    n = (16, 12)[:nd]
    g = O.Grid((0.0,) * nd, (1.0,) * nd, n, vertex_funcs=[stretched(m, 0.05 + 0.02 * d) for d, m in enumerate(n)])
    fn = (lambda x: 0.3 + 1.7 * x) if nd == 1 else (lambda x, y: 0.3 + 1.7 * x - 0.6 * y + 2.1 * x * y)
    for src in _locs(nd):
        f = O.HField(g, src)
        xs = [g.coords(src, d, first=1 - 2, count=f.n[d] + 4) for d in range(nd)]
        f.xyz()[...] = fn(*np.meshgrid(*xs, indexing="ij"))
        for to in _locs(nd):
            h = O.HField(g, to)
            O.op("lerp", g, h, f, b=O.interior_box(h))
            xt = [g.coords(to, d) for d in range(nd)]
            want = fn(*np.meshgrid(*xt, indexing="ij"))
            got = h.xyz()[tuple(slice(2, 2 + m) for m in h.n)]
            assert np.max(np.abs(got - want)) <= 8 * np.finfo(float).eps * np.max(np.abs(want)), (src, to)


def _spacings(fn, i):
    v = lambda k: float(fn(k))
    c = lambda k: 0.5 * (v(k) + v(k + 1))
    return (v(i + 1) - v(i)), (c(i) - c(i - 1))    # Δ(Center, i), Δ(Vertex, i)


def test_partial_scales_by_the_source_location_spacing():
    """∂ = δ·iΔ(grid, loc(f), dim, I) (partial_derivatives.jl:5,
    field_operators.jl:20-24) on a stretched 1D axis, for a Center and a
    Vertex field."""
    n = 12
    fn = stretched(n)
    g = O.Grid((0.0,), (1.0,), (n,), vertex_funcs=[fn])
    rng = np.random.default_rng(3)
    for loc in (C_, V_):
        f = _rand(g, (loc,), rng)
        to = (1 - loc,)
        a, d = O.HField(g, to), O.HField(g, to)
        O.op("partial", g, a, f)
        O.op("delta", g, d, f)
        for I in range(0, n + 2):
            dc, dv = _spacings(fn, I)
            want = d.arr[I + 1] * (1.0 / (dc if loc == C_ else dv))
            assert a.arr[I + 1] == want, (loc, I)


def test_boundary_ghosts_use_the_reference_indices():
    """Center Dirichlet: f[I_f] = muladd(2 Δ(Center, I_f2) iΔ(Vertex, I_f), a - b, b);
    Neumann: f[I_f] = muladd(Δ(flip(loc), I_f), ∓q, f[I_f2])
    (first_order_boundary_condition.jl:59-84) on a stretched 1D axis."""
    n = 10
    fn = stretched(n, 0.08)
    g = O.Grid((0.0,), (1.0,), (n,), vertex_funcs=[fn])
    rng = np.random.default_rng(4)
    a_val, q_val = 0.7, -0.3
    for loc in (C_, V_):
        f = _rand(g, (loc,), rng)
        m = f.n[0]
        for side, (ig, inn) in ((1, (0, 1)), (2, (m + 1, m))):
            h = O.HField(g, (loc,))
            h.arr[...] = f.arr
            O.bc_apply(g, [(h, {0: (O.Neumann(q_val), O.Neumann(q_val))})])
            dc, dv = _spacings(fn, ig)
            sp = dv if loc == C_ else dc                 # Δ(flip(loc), I_f)
            qs = -q_val if side == 1 else q_val
            assert h.arr[ig + 1] == fma(sp, qs, f.arr[inn + 1]), (loc, side)
            if loc == C_:
                h.arr[...] = f.arr
                O.bc_apply(g, [(h, {0: (O.Dirichlet(a_val), O.Dirichlet(a_val))})])
                dc2, _ = _spacings(fn, inn)
                _, dv_g = _spacings(fn, ig)
                t = (2.0 * dc2) * (1.0 / dv_g)
                b = f.arr[inn + 1]
                assert h.arr[ig + 1] == fma(t, a_val - b, b), side
