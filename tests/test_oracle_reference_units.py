"""Pins the CPU oracle against the reference's own unit-test assertions.

Each test mirrors one assertion block of Chmy.jl's test suite (cited per
test); no golden files exist in the reference, so these closed forms and
operator identities are the reference's own known answers.
"""
import math

import numpy as np
import pytest

import oracle as O
from oracle import CENTER as C, VERTEX as V

TYPES = [np.float64, np.float32]


# ---------------------------------------------------------------- grids
def test_grid_metrics_and_coords():
    """test/test_grids.jl:10-175 on 5x20 over [-1,1]x[-2,2]."""
    g = O.Grid((-1, -2), (2, 4), (5, 20))
    assert g.size(C) == (5, 20) and g.size(V) == (6, 21)
    assert g.size((C, V)) == (5, 21) and g.size((V, C)) == (6, 20)
    assert g.spacing == (0.4, 0.2)
    assert g.inv_spacing == (2.5, 5.0)
    xv, yv = g.coords(V, 0), g.coords(V, 1)
    xc, yc = g.coords(C, 0), g.coords(C, 1)
    # bounds (test_grids.jl "bounds")
    assert np.isclose(xv[0], -1.0) and np.isclose(xv[-1], 1.0)
    assert np.isclose(yv[0], -2.0) and np.isclose(yv[-1], 2.0)
    assert np.isclose(xc[0], -0.8) and np.isclose(xc[-1], 0.8)
    assert np.isclose(yc[0], -1.9) and np.isclose(yc[-1], 1.9)
    # extent (Center extent = extent - spacing)
    assert np.isclose(xc[-1] - xc[0], 1.6) and np.isclose(yc[-1] - yc[0], 3.8)
    # fma form of coord: vertex(i) = fma(i-1, dx, o)
    for i in range(1, 7):
        assert xv[i - 1] == math.fsum([(i - 1) * 0.4, -1.0]) or np.isclose(xv[i - 1], -1 + (i - 1) * 0.4)


def test_field_coords_tuple_halo():
    """test/test_fields.jl:19-55: 2x2x2 grid on [0,1]^3, loc (C,V,C), halo (1,0,1)."""
    g = O.Grid((0, 0, 0), (1, 1, 1), (2, 2, 2))
    f = O.HField(g, (C, V, C), halo=(1, 0, 1))
    assert f.n == (2, 3, 2) and f.pad == (6, 3, 6)
    y = g.coords((C, V, C), 1)
    x = g.coords((C, V, C), 0)
    assert list(y) == [0.0, 0.5, 1.0]
    assert list(x) == [0.25, 0.75]
    f.set_interior(np.broadcast_to(y[None, :, None], f.n))
    exp = np.array([0.0, 0.0, 0.5, 0.5, 1.0, 1.0, 0.0, 0.0, 0.5, 0.5, 1.0, 1.0])
    assert np.array_equal(f.interior().reshape(-1, order="F"), exp)


# ---------------------------------------------------------------- interpolation
def _av4(A):
    return 0.25 * (A[:-1, :-1] + A[1:, :-1] + A[1:, 1:] + A[:-1, 1:])


def _avx(A):
    return 0.5 * (A[:-1, :] + A[1:, :])


def _avy(A):
    return 0.5 * (A[:, :-1] + A[:, 1:])


@pytest.mark.parametrize("T", TYPES)
def test_lerp_center_source(T):
    """test/test_interpolations.jl:14-45 (c2v, c2c, c2cv, c2vc)."""
    g = O.Grid((0, 0), (1, 1), (2, 2))
    fc = O.HField(g, C, T)
    fc.set_interior(np.arange(1, 5).reshape(2, 2, order="F"))
    fci = fc.interior().astype(np.float64)
    out = {}
    for name, loc in (("c2v", V), ("c2c", C), ("c2cv", (C, V)), ("c2vc", (V, C))):
        d = O.HField(g, loc, T)
        O.op("lerp", g, d, fc, b=O.interior_box(d))
        out[name] = d.interior().astype(np.float64)
    assert np.allclose(out["c2v"][1:-1, 1:-1], _av4(fci))
    assert np.allclose(out["c2c"], fci)
    assert np.allclose(out["c2cv"][:, 1:-1], _avy(fci))
    assert np.allclose(out["c2vc"][1:-1, :], _avx(fci))


@pytest.mark.parametrize("T", TYPES)
def test_lerp_vertex_source(T):
    """test/test_interpolations.jl:46-74 (v2c, v2v, v2cv, v2vc)."""
    g = O.Grid((0, 0), (1, 1), (2, 2))
    fv = O.HField(g, V, T)
    fv.set_interior(np.arange(1, 10).reshape(3, 3, order="F"))
    fvi = fv.interior().astype(np.float64)
    out = {}
    for name, loc in (("v2c", C), ("v2v", V), ("v2cv", (C, V)), ("v2vc", (V, C))):
        d = O.HField(g, loc, T)
        O.op("lerp", g, d, fv, b=O.interior_box(d))
        out[name] = d.interior().astype(np.float64)
    assert np.allclose(out["v2c"], _av4(fvi))
    assert np.allclose(out["v2v"], fvi)
    assert np.allclose(out["v2cv"], _avx(fvi))
    assert np.allclose(out["v2vc"], _avy(fvi))


# ---------------------------------------------------------------- grid operators
def _gauss3(g, f):
    x = g.coords(f.loc, 0)[:, None, None]
    y = g.coords(f.loc, 1)[None, :, None]
    z = g.coords(f.loc, 2)[None, None, :]
    f.set_interior(np.exp(-x ** 2 - y ** 2 - z ** 2))


@pytest.mark.parametrize("T", TYPES)
def test_divg_equals_sum_of_partials(T):
    """test/test_grid_operators.jl:14-42: divg(V) == ∂xVx + ∂yVy + ∂zVz (exact)."""
    g = O.Grid((-5, -5, -5), (10, 10, 10), (12, 10, 8))
    Ci = O.HField(g, C, T)
    _gauss3(g, Ci)
    Vf = [O.HField(g, tuple(V if d == D else C for d in range(3)), T) for D in range(3)]
    for D in range(3):
        O.op("partial", g, Vf[D], Ci, dim=D)
    C1 = O.HField(g, C, T)
    C2 = O.HField(g, C, T)
    parts = [O.HField(g, C, T) for _ in range(3)]
    for D in range(3):
        O.op("partial", g, parts[D], Vf[D], dim=D)
    C1.arr[...] = (parts[0].arr + parts[1].arr) + parts[2].arr
    O.op("divg", g, C2, Vf)
    assert np.array_equal(C2.interior(), C1.interior())


@pytest.mark.parametrize("T", TYPES)
def test_lapl_equals_sum_of_second_partials(T):
    """test/test_grid_operators.jl:44-62: lapl == Σ∂² (exact)."""
    g = O.Grid((-5, -5, -5), (10, 10, 10), (12, 10, 8))
    Ci = O.HField(g, C, T)
    _gauss3(g, Ci)
    parts = [O.HField(g, C, T) for _ in range(3)]
    for D in range(3):
        O.op("d2", g, parts[D], Ci, dim=D)
    C2 = O.HField(g, C, T)
    O.op("lapl", g, C2, Ci)
    assert np.array_equal(C2.interior(), ((parts[0].arr + parts[1].arr) + parts[2].arr).T[2:-2, 2:-2, 2:-2])


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("chi_loc", [C, V])
def test_divg_grad(T, chi_loc):
    """test/test_grid_operators.jl:64-112: divg_grad(C, χ) ≈ Σ∂(lerp(χ)∂C)."""
    g = O.Grid((-5, -5, -5), (10, 10, 10), (12, 10, 8))
    Ci = O.HField(g, C, T)
    _gauss3(g, Ci)
    chi = O.HField(g, chi_loc, T)
    _gauss3(g, chi)
    Vf = []
    for D in range(3):
        loc = tuple(V if d == D else C for d in range(3))
        a = O.HField(g, loc, T)
        b = O.HField(g, loc, T)
        O.op("lerp", g, a, chi)
        O.op("partial", g, b, Ci, dim=D)
        a.arr[...] = a.arr * b.arr
        Vf.append(a)
    C1 = O.HField(g, C, T)
    O.op("divg", g, C1, Vf)
    C2 = O.HField(g, C, T)
    O.op("divg_grad", g, C2, Ci, k=chi)
    assert np.allclose(C2.interior(), C1.interior(), rtol=1e-5 if T == np.float32 else 1e-12, atol=1e-6)


@pytest.mark.parametrize("T", TYPES)
def test_vmag_constant(T):
    """test/test_grid_operators.jl:114-132: vmag of a constant 2-vector in 3D ≈ 3.4641."""
    g = O.Grid((-5, -5, -5), (10, 10, 10), (12, 10, 8))
    Vf = []
    for D in range(3):
        f = O.HField(g, tuple(V if d == D else C for d in range(3)), T)
        f.set_interior(2.0)
        Vf.append(f)
    out = O.HField(g, C, T)
    O.op("vmag", g, out, Vf, b=O.interior_box(out))
    assert np.all(np.round(out.interior().astype(np.float64), 4) == 3.4641)


# ---------------------------------------------------------------- boundary conditions
@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("loc", [C, V])
def test_bc_1d(T, loc):
    """test/test_boundary_conditions.jl:11-88 (1D Center and Vertex, n=8 on [-π,π])."""
    g = O.Grid((-math.pi,), (2 * math.pi,), (8,))
    f = O.HField(g, loc, T)
    dx = g.spacing[0]
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Dirichlet())])
    fi = f.interior(with_halo=True).astype(np.float64)
    if loc == C:
        assert np.isclose(fi[0], -fi[1]) and np.isclose(fi[-1], -fi[-2])
    else:
        assert fi[1] == 0.0 and fi[-2] == 0.0
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Neumann())])
    fi = f.interior(with_halo=True)
    assert fi[0] == fi[1] and fi[-1] == fi[-2]
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Dirichlet(2.0))])
    fi = f.interior(with_halo=True).astype(np.float64)
    if loc == C:
        assert np.isclose(fi[0], -fi[1] + 4.0) and np.isclose(fi[-1], -fi[-2] + 4.0)
    else:
        assert fi[1] == 2.0 and fi[-2] == 2.0
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Neumann(2.0))])
    fi = f.interior(with_halo=True).astype(np.float64)
    tol = 1e-5 if T == np.float32 else 1e-12
    assert np.isclose((fi[1] - fi[0]) / dx, 2.0, rtol=tol)
    assert np.isclose((fi[-1] - fi[-2]) / dx, 2.0, rtol=tol)


@pytest.mark.parametrize("T", TYPES)
def test_bc_2d(T):
    """test/test_boundary_conditions.jl:90-138 (2D, loc (Center, Vertex), 8x8 on [-π,π]²)."""
    g = O.Grid((-math.pi, -math.pi), (2 * math.pi, 2 * math.pi), (8, 8))
    f = O.HField(g, (C, V), T)
    dx, dy = g.spacing
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Dirichlet())])
    fi = f.interior(with_halo=True).astype(np.float64)
    assert np.allclose(fi[0, 1:-1], -fi[1, 1:-1]) and np.allclose(fi[-1, 1:-1], -fi[-2, 1:-1])
    assert np.all(fi[1:-1, 1] == 0) and np.all(fi[1:-1, -2] == 0)
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Neumann())])
    fi = f.interior(with_halo=True)
    assert np.all(fi[0, 1:-1] == fi[1, 1:-1]) and np.all(fi[-1, 1:-1] == fi[-2, 1:-1])
    assert np.all(fi[1:-1, 0] == fi[1:-1, 1]) and np.all(fi[1:-1, -1] == fi[1:-1, -2])
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Dirichlet(2.0))])
    fi = f.interior(with_halo=True).astype(np.float64)
    assert np.allclose(fi[0, 1:-1], -fi[1, 1:-1] + 4) and np.allclose(fi[-1, 1:-1], -fi[-2, 1:-1] + 4)
    assert np.all(fi[1:-1, 1] == 2) and np.all(fi[1:-1, -2] == 2)
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Neumann(2.0))])
    fi = f.interior(with_halo=True).astype(np.float64)
    tol = 1e-5 if T == np.float32 else 1e-12
    assert np.allclose((fi[1, 1:-1] - fi[0, 1:-1]) / dx, 2.0, rtol=tol)
    assert np.allclose((fi[-1, 1:-1] - fi[-2, 1:-1]) / dx, 2.0, rtol=tol)
    assert np.allclose((fi[1:-1, 1] - fi[1:-1, 0]) / dy, 2.0, rtol=tol)
    assert np.allclose((fi[1:-1, -1] - fi[1:-1, -2]) / dy, 2.0, rtol=tol)


@pytest.mark.parametrize("T", TYPES)
def test_bc_3d(T):
    """test/test_boundary_conditions.jl:140-205 (3D, loc (C,V,C), 8x8x6 on [-π,π]³)."""
    g = O.Grid((-math.pi,) * 3, (2 * math.pi,) * 3, (8, 8, 6))
    f = O.HField(g, (C, V, C), T)
    dx, dy, dz = g.spacing
    s = slice(1, -1)
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Dirichlet())])
    fi = f.interior(with_halo=True).astype(np.float64)
    assert np.allclose(fi[0, s, s], -fi[1, s, s]) and np.allclose(fi[-1, s, s], -fi[-2, s, s])
    assert np.all(fi[s, 1, s] == 0) and np.all(fi[s, -2, s] == 0)
    assert np.allclose(fi[s, s, 0], -fi[s, s, 1]) and np.allclose(fi[s, s, -1], -fi[s, s, -2])
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Neumann())])
    fi = f.interior(with_halo=True)
    assert np.all(fi[0, s, s] == fi[1, s, s]) and np.all(fi[-1, s, s] == fi[-2, s, s])
    assert np.all(fi[s, 0, s] == fi[s, 1, s]) and np.all(fi[s, -1, s] == fi[s, -2, s])
    assert np.all(fi[s, s, 0] == fi[s, s, 1]) and np.all(fi[s, s, -1] == fi[s, s, -2])
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Dirichlet(2.0))])
    fi = f.interior(with_halo=True).astype(np.float64)
    assert np.allclose(fi[0, s, s], -fi[1, s, s] + 4) and np.allclose(fi[-1, s, s], -fi[-2, s, s] + 4)
    assert np.all(fi[s, 1, s] == 2) and np.all(fi[s, -2, s] == 2)
    assert np.allclose(fi[s, s, 0], -fi[s, s, 1] + 4) and np.allclose(fi[s, s, -1], -fi[s, s, -2] + 4)
    f.set_interior(1)
    O.bc_apply(g, [(f, O.Neumann(2.0))])
    fi = f.interior(with_halo=True).astype(np.float64)
    tol = 1e-5 if T == np.float32 else 1e-12
    assert np.allclose((fi[1, s, s] - fi[0, s, s]) / dx, 2.0, rtol=tol)
    assert np.allclose((fi[-1, s, s] - fi[-2, s, s]) / dx, 2.0, rtol=tol)
    assert np.allclose((fi[s, 1, s] - fi[s, 0, s]) / dy, 2.0, rtol=tol)
    assert np.allclose((fi[s, -1, s] - fi[s, -2, s]) / dy, 2.0, rtol=tol)
    assert np.allclose((fi[s, s, 1] - fi[s, s, 0]) / dz, 2.0, rtol=tol)
    assert np.allclose((fi[s, s, -1] - fi[s, s, -2]) / dz, 2.0, rtol=tol)
