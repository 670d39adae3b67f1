"""Boundary conditions whose value is a BoundaryFunction
(src/BoundaryConditions/boundary_function.jl): the ghost relations of
first_order_boundary_condition.jl hold with the function evaluated where the
reference evaluates it, for the continuous (reduced and full dimensions) and
discrete forms, on random storage; and the device batch equals the oracle's
with the same value field."""
import numpy as np
import pytest

import chmy
from chmy import Center as Cc, Vertex as Vv
import oracle as O
from _helpers import dev, ogrid, pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def test_center_dirichlet_reduced_dims_matches_reference_example(arch):
    """The docstring example of boundary_function.jl:60-70: on a Center field,
    x = Dirichlet(BoundaryFunction((y, ly) -> y (ly - y); parameters=(2,))),
    y = Neumann(): the wall value (f[0] + f[1]) / 2 equals the function at the
    ghost's tangential Center coordinate, on both sides, over the tangential
    range 0..n+2; the batch equals the oracle's with the same value field."""
    g = chmy.UniformGrid(arch, origin=(0.0, 0.0), extent=(2.0, 2.0), dims=(10, 12))
    og = ogrid(g)
    rng = np.random.default_rng(3)
    f, hf = pair(arch, g, og, Cc(), np.float64, rng=rng)
    fun = lambda y, ly: y * (ly - y)
    bs = chmy.bc_(arch, g, (f, {"x": chmy.Dirichlet(chmy.BoundaryFunction(fun, parameters=(2.0,))),
                                "y": chmy.Neumann()}))
    chmy.synchronize()
    a = dev(f)                                     # storage (y, x), logical i at i + 1
    nx, ny = g.dims
    ys = g.coords(Cc(), 1, first=0, count=ny + 3)   # logical 0 .. n+2
    for side, (gi, ii) in enumerate(((0, 1), (nx + 1, nx))):
        wall = 0.5 * (a[1:ny + 4, gi + 1] + a[1:ny + 4, ii + 1])
        np.testing.assert_allclose(wall, fun(ys, 2.0), rtol=1e-13, atol=1e-13, err_msg=f"side {side + 1}")
    # the oracle with the same value field (x side 1) gives the same bits
    vals = [bs[0][s].conditions[0].value for s in range(2)]
    tg = vals[0].grid
    otg = ogrid(tg)
    hv = []
    for v in vals:
        h = O.HField(otg, O.CENTER)
        h.arr[...] = dev(v)
        hv.append(h)
    O.bc_apply(og, [(hf, {0: (O.Dirichlet(hv[0]), O.Dirichlet(hv[1])), 1: O.Neumann()})])
    assert np.array_equal(a, hf.arr)


def test_vertex_dirichlet_full_dims_3d(arch):
    """A Vertex-located Dirichlet sets the boundary node itself to the function
    of all three coordinates of that node (f's own location), exactly."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.25, 0.0), extent=(1.0, 1.5, 0.75), dims=(9, 7, 5))
    og = ogrid(g)
    rng = np.random.default_rng(4)
    loc = (Vv(), Cc(), Cc())
    f, _ = pair(arch, g, og, loc, np.float64, rng=rng)
    fun = lambda x, y, z: x + 2.0 * y - z * z
    chmy.bc_(arch, g, (f, {"x": chmy.Dirichlet(chmy.BoundaryFunction(fun, reduce_dims=False))}))
    chmy.synchronize()
    a = dev(f)                                     # (z, y, x)
    nx, ny, nz = g.dims
    ys = g.coords(Cc(), 1, first=0, count=ny + 3)
    zs = g.coords(Cc(), 2, first=0, count=nz + 3)
    Z, Y = np.meshgrid(zs, ys, indexing="ij")
    for node in (1, nx + 1):
        x = g.coords(Vv(), 0, first=node, count=1)[0]
        want = np.vectorize(fun)(x, Y, Z)
        assert np.array_equal(a[1:nz + 4, 1:ny + 4, node + 1], want), node


def test_neumann_discrete_on_vertex_field(arch):
    """A discrete BoundaryFunction (grid, loc, dim, i) -> q on a y-Vertex field
    with Neumann on y: f[m+1] - f[m] = Δy q and f[1] - f[0] = Δy q at each
    tangential index i, the function receiving Chmy's logical index; the
    location it is given is the field's, flipped."""
    g = chmy.UniformGrid(arch, origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(11, 8))
    og = ogrid(g)
    rng = np.random.default_rng(5)
    f, _ = pair(arch, g, og, (Cc(), Vv()), np.float64, rng=rng)
    seen = set()

    def q(grid, loc, dim, i):
        seen.add((tuple(type(l).__name__ for l in loc), dim))
        return 0.25 * i - 1.0

    chmy.bc_(arch, g, (f, {"y": chmy.Neumann(chmy.BoundaryFunction(q, discrete=True))}))
    chmy.synchronize()
    assert seen == {(("Vertex", "Center"), 1)}
    a = dev(f)
    nx, ny = g.dims
    m = ny + 1
    h = chmy.spacing(g)[1]
    i = np.arange(0, nx + 3)
    want = 0.25 * i - 1.0
    np.testing.assert_allclose(a[m + 2, 1:nx + 4] - a[m + 1, 1:nx + 4], h * want, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(a[2, 1:nx + 4] - a[1, 1:nx + 4], h * want, rtol=1e-12, atol=1e-14)


def test_plain_function_and_1d(arch):
    """A plain function is a continuous, reduced BoundaryFunction
    (boundary_function.jl:78-81); in 1D it reduces to a number per side."""
    g = chmy.UniformGrid(arch, origin=(0.0,), extent=(1.0,), dims=(16,))
    og = ogrid(g)
    rng = np.random.default_rng(6)
    f, _ = pair(arch, g, og, Cc(), np.float64, rng=rng)
    chmy.bc_(arch, g, (f, chmy.Dirichlet(lambda: 3.5)))
    chmy.synchronize()
    a = dev(f)
    n = g.dims[0]
    np.testing.assert_allclose([0.5 * (a[1] + a[2]), 0.5 * (a[n + 2] + a[n + 1])], [3.5, 3.5], rtol=1e-15)
