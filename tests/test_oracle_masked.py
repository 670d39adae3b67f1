"""Pins of the oracle's masked operators (src/GridOperators/masked_operators.jl,
masked_field_operators.jl, field_mask.jl) against the unmasked operators,
which the reference's own tests and the closed forms pin.

By the reference's definition left(f, ω)[I] = f[Il]·at(ω, loc(f), Il), so a
masked operator on f equals the unmasked operator on the elementwise product
f⊙ω_loc(f), where ω_loc(f) is the mask component at f's location: exactly,
since each knot's product is the same IEEE multiplication either way.  With
ω ≡ 1 the masked operators are the unmasked ones.  The first identity fixes
where the mask is read (the knot, at the source's location); reading it at
the output index, a plausible mistake planted in the oracle, fails it."""
import numpy as np
import pytest

import oracle as O

C_, V_ = O.CENTER, O.VERTEX


def _grid(nd):
    dims = [(11,), (9, 7), (7, 6, 5)][nd - 1]
    return O.Grid((-0.3,) * nd, (1.3,) * nd, dims)


def _rand(g, loc, rng, lo=-1.0, hi=1.0):
    f = O.HField(g, loc)
    f.arr[...] = rng.uniform(lo, hi, f.arr.shape)
    return f


def _mask(g, loc, rng, kind):
    w = O.HField(g, loc)
    if kind == "ones":
        w.arr[...] = 1.0
    elif kind == "indicator":
        w.arr[...] = (rng.random(w.arr.shape) < 0.7).astype(np.float64)
    else:
        w.arr[...] = rng.random(w.arr.shape)
    return w


def _product(f, w):
    p = O.HField(f.grid, f.loc)
    p.arr[...] = f.arr * w.arr
    return p


def _locs(nd):
    out = []
    for q in range(2 ** nd):
        out.append(tuple((q >> d) & 1 for d in range(nd)))
    return out


def _flip(loc, D):
    l = list(loc)
    l[D] = 1 - l[D]
    return tuple(l)


CASES = []
for nd in (1, 2, 3):
    for loc in _locs(nd):
        for D in range(nd):
            CASES.append((nd, loc, D))


def _check_single(nd, loc, D, kind, rng, names=("left", "right", "delta", "partial", "d2", "dkd")):
    g = _grid(nd)
    f = _rand(g, loc, rng)
    w = _mask(g, loc, rng, kind)
    k = _rand(g, loc, rng, 0.5, 2.0)
    fw = f if kind == "ones" else _product(f, w)
    for name in names:
        to = loc if name in ("d2", "dkd") else _flip(loc, D)
        a, b = O.HField(g, to), O.HField(g, to)
        box = O.interior_box(a)
        O.op_masked(name, g, a, f, w, k=k if name == "dkd" else None, dim=D, b=box)
        O.op(name, g, b, fw, k=k if name == "dkd" else None, dim=D, b=box)
        if not np.array_equal(a.arr, b.arr):
            return name
    return None


@pytest.mark.parametrize("nd,loc,D", CASES)
@pytest.mark.parametrize("kind", ["ones", "indicator", "fraction"])
def test_masked_operators_are_unmasked_operators_of_the_masked_field(nd, loc, D, kind):
    rng = np.random.default_rng(100 * nd + 10 * D + sum(loc))
    assert _check_single(nd, loc, D, kind, rng) is None


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("kind", ["ones", "fraction"])
def test_masked_sums_divg_lapl_divg_grad(nd, kind):
    """divg(V, ω) with each component's mask at its own location; lapl and
    divg_grad of a Center field: the unmasked sums of the masked fields."""
    rng = np.random.default_rng(7 + nd)
    g = _grid(nd)
    vl = [tuple(1 if d == D else 0 for d in range(nd)) for D in range(nd)]
    V = [_rand(g, l, rng) for l in vl]
    W = [_mask(g, l, rng, kind) for l in vl]
    a, b = O.HField(g, C_), O.HField(g, C_)
    box = O.interior_box(a)
    O.op_masked("divg", g, a, V, W, b=box)
    O.op("divg", g, b, [_product(v, w) for v, w in zip(V, W)], b=box)
    assert np.array_equal(a.arr, b.arr)
    F, w, K = _rand(g, C_, rng), _mask(g, C_, rng, kind), _rand(g, C_, rng, 0.5, 2.0)
    for name in ("lapl", "divg_grad"):
        a, b = O.HField(g, C_), O.HField(g, C_)
        kk = K if name == "divg_grad" else None
        O.op_masked(name, g, a, F, w, k=kk, b=box)
        O.op(name, g, b, _product(F, w), k=kk, b=box)
        assert np.array_equal(a.arr, b.arr), name


def test_a_full_mask_on_a_sub_box_is_the_unmasked_operator():
    """ω = 1 inside a region, 0 outside: away from the region's edge the
    masked ∂ equals the unmasked ∂ of f, and outside it is 0."""
    g = _grid(2)
    rng = np.random.default_rng(3)
    f = _rand(g, C_, rng)
    w = O.HField(g, C_)
    w.xyz()[3:8, 2:7] = 1.0            # storage indices: logical cells 2..6 x 1..5
    a, b = O.HField(g, (V_, C_)), O.HField(g, (V_, C_))
    O.op_masked("partial", g, a, f, w, dim=0, b=O.interior_box(a))
    O.op("partial", g, b, f, dim=0, b=O.interior_box(b))
    # vertex i reads cells i-1, i: both inside for i = 3..6 (storage 4..7), y cells 1..5 (storage 2..6)
    assert np.array_equal(a.xyz()[4:8, 2:7], b.xyz()[4:8, 2:7])
    assert not a.xyz()[2:3, 2:8].any() and not a.xyz()[9:11, 2:8].any()


def test_planted_mask_index_fails():
    rng = np.random.default_rng(5)
    with O.planted("mask_at_out"):
        bad = [c for c in CASES if _check_single(*c, "fraction", rng, names=("left", "right")) is not None]
    assert len(bad) >= len(CASES) // 2
