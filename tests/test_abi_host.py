"""CPU-side checks: the C-ABI library loads and exports every declared entry
point; descriptor validation rejects misuse without touching a GPU; host
logic (grid metrics, coordinates, batch regularisation, Launcher boxes)
matches the reference's definitions."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest
import torch

import chmy
from chmy import _lib
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chmy_mi355x.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(chmy_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers all of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_validation_rejects_misuse_on_host():
    """Descriptor checks run before any HIP call: wrong locations, sizes,
    halos and boxes return negative status with a message."""
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(8, 6))
    gd = g.desc()

    def fdesc(loc, n, halo=1, dtype=_lib.F64):
        f = _lib.ChmyField()
        f.data = 0x1000  # never dereferenced: validation fails first
        f.nd = 2
        f.dtype = dtype
        for d in range(2):
            f.n[d] = n[d]
            f.halo[d] = halo
            f.loc[d] = loc[d]
        return f

    Cf = fdesc((0, 0), (8, 6))
    q_bad = (_lib.ChmyField * 2)(fdesc((0, 0), (8, 6)), fdesc((0, 1), (8, 7)))  # q.x at Center: wrong
    rc = _lib.lib().chmy_diffusion_flux(C.byref(gd), None, q_bad, C.byref(Cf), 1.0, None)
    assert rc == -3 and b"location" in _lib.lib().chmy_last_error()
    q_ok = (_lib.ChmyField * 2)(fdesc((1, 0), (9, 6)), fdesc((0, 1), (8, 7)))
    box = _lib.ChmyBox()
    box.lo[0], box.lo[1], box.len[0], box.len[1] = 0, 0, 11, 8  # x range [0, 11) > n+2 = 10
    rc = _lib.lib().chmy_diffusion_flux(C.byref(gd), C.byref(box), q_ok, C.byref(Cf), 1.0, None)
    assert rc == -4 and b"box" in _lib.lib().chmy_last_error()
    Ch = fdesc((0, 0), (8, 6), halo=0)
    rc = _lib.lib().chmy_diffusion_flux(C.byref(gd), None, q_ok, C.byref(Ch), 1.0, None)
    assert rc == -5
    Cs = fdesc((0, 0), (8, 5))
    rc = _lib.lib().chmy_diffusion_flux(C.byref(gd), None, q_ok, C.byref(Cs), 1.0, None)
    assert rc == -3 and b"size" in _lib.lib().chmy_last_error()
    q32 = (_lib.ChmyField * 2)(fdesc((1, 0), (9, 6), dtype=_lib.F32), fdesc((0, 1), (8, 7), dtype=_lib.F32))
    rc = _lib.lib().chmy_diffusion_flux(C.byref(gd), None, q32, C.byref(Cf), 1.0, None)
    assert rc == -2


def test_runtime_helpers_validate_on_host():
    """Stream and fill helpers reject misuse before any HIP call; the fused
    diffusion entry points check their buffers like the per-kernel ones."""
    L = _lib.lib()
    s = C.c_void_p()
    assert L.chmy_stream_create(1, C.byref(s)) == -1 and b"priority" in L.chmy_last_error()
    assert L.chmy_stream_create(0, None) == -1
    assert L.chmy_stream_destroy(None) == -1
    assert L.chmy_memset_zero(None, 16, None) == -1
    assert L.chmy_memset_zero(None, 0, None) == 0        # nothing to do
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(8, 6))
    f = _lib.ChmyField()
    f.data, f.nd, f.dtype = 0x1000, 2, _lib.F64
    for d, n in enumerate((8, 6)):
        f.n[d], f.halo[d], f.loc[d] = n, 1, 0
    q = (_lib.ChmyField * 2)()
    for D in range(2):
        q[D].data, q[D].nd, q[D].dtype = 0x2000, 2, _lib.F64
        for d, n in enumerate((8, 6)):
            q[D].n[d], q[D].halo[d], q[D].loc[d] = n + (d == D), 1, int(d == D)
    rc = L.chmy_diffusion_step(C.byref(g.desc()), None, C.byref(f), C.byref(f), q, 1.0, 0.1, None)
    assert rc == -1 and b"distinct" in L.chmy_last_error()
    par = C.c_int(2)
    counts = (C.c_int32 * 6)()
    rc = L.chmy_diffusion_iterate(C.byref(g.desc()), C.byref(f), C.byref(f), q, counts, None, 1.0, 0.1, 1,
                                  C.byref(par), None)
    assert rc == -1 and b"parity" in L.chmy_last_error()


def test_adjoint_entry_points_validate_on_host():
    """SIA run/reverse and the Stokes gradient pass reject misuse before any
    launch: missing arrays, aliasing, checkpoint counts, wrong locations."""
    L = _lib.lib()
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(8, 6))
    gd = g.desc()

    def fdesc(loc, n, data=0x1000):
        f = _lib.ChmyField()
        f.data = data
        f.nd = 2
        f.dtype = _lib.F64
        for d in range(2):
            f.n[d], f.halo[d], f.loc[d] = n[d], 1, loc[d]
        return f

    Cc = lambda data=0x1000: fdesc((0, 0), (8, 6), data)
    dts = (C.c_double * 4)(1.0, 1.0, 1.0, 1.0)
    B = Cc(0x9000)
    # states alias: states[0] and states[1] share storage
    st = (_lib.ChmyField * 3)(Cc(0x1000), Cc(0x1000), Cc(0x2000))
    assert L.chmy_sia_run(C.byref(gd), st, 3, C.byref(B), None, 0.0, 3, 0.0, 0.0, 0.0, dts, None) == -1
    assert b"share storage" in L.chmy_last_error()
    # a state at the wrong location
    st = (_lib.ChmyField * 2)(Cc(0x1000), fdesc((1, 0), (9, 6), 0x2000))
    assert L.chmy_sia_run(C.byref(gd), st, 2, C.byref(B), None, 0.0, 3, 0.0, 0.0, 0.0, dts, None) == -3
    # reverse: checkpoint count must be ceil(nsteps / stride)
    ck = (_lib.ChmyField * 2)(Cc(0x1000), Cc(0x2000))
    seg = (_lib.ChmyField * 2)(Cc(0x3000), Cc(0x4000))
    la, lb = Cc(0x5000), Cc(0x6000)
    slot = C.c_int(-1)
    rc = L.chmy_sia_reverse(C.byref(gd), ck, 2, 2, 5, seg, C.byref(la), C.byref(lb), None, C.byref(B), None, 0.0,
                            3, 0.0, 0.0, 0.0, dts, C.byref(slot), None)
    assert rc == -1 and b"nckpt" in L.chmy_last_error()
    rc = L.chmy_sia_reverse(C.byref(gd), ck, 2, 2, 4, seg, C.byref(la), C.byref(la), None, C.byref(B), None, 0.0,
                            3, 0.0, 0.0, 0.0, dts, C.byref(slot), None)
    assert rc == -1 and b"lam_a and lam_b" in L.chmy_last_error()
    # the no-allocation variant: same validation, and a workspace above its
    # 48 MiB device scratch is refused before any HIP call
    rc = L.chmy_sia_reverse_noalloc(C.byref(gd), ck, 2, 2, 5, C.byref(la), None, C.byref(B), None, 0.0, 3, 0.0,
                                    0.0, 0.0, dts, None)
    assert rc == -1 and b"nckpt" in L.chmy_last_error()
    gbig = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(1024, 1024)).desc()
    Cb = lambda data: fdesc((0, 0), (1024, 1024), data)
    ckb = (_lib.ChmyField * 30)(*[Cb(0x10000 * (q + 1)) for q in range(30)])
    Bb, lb_ = Cb(0x400000), Cb(0x500000)
    dtb = (C.c_double * 300)(*([1.0] * 300))
    rc = L.chmy_sia_reverse_noalloc(C.byref(gbig), ckb, 30, 10, 300, C.byref(lb_), None, C.byref(Bb), None, 0.0,
                                    3, 0.0, 0.0, 0.0, dtb, None)
    assert rc == -7 and b"scratch" in L.chmy_last_error()
    # Stokes gradient pass: tau_a with a wrong location
    tl = [(0, 0), (0, 0), (1, 1)]
    tn = lambda l: (8 + l[0], 6 + l[1])
    tau = (_lib.ChmyField * 3)(*[fdesc(l, tn(l)) for l in tl])
    bad = (_lib.ChmyField * 3)(*[fdesc(l, tn(l)) for l in tl[:2]], fdesc((1, 0), (9, 6)))
    eta, gr = Cc(), Cc()
    assert L.chmy_stokes_grad_eta(C.byref(gd), tau, bad, C.byref(eta), C.byref(gr), 1, None) == -3
    assert L.chmy_stokes_grad_eta(C.byref(gd), tau, None, C.byref(eta), C.byref(gr), 1, None) == -1


def test_host_field_refuses_kernel_use():
    g = chmy.UniformGrid(origin=(0.0,), extent=(1.0,), dims=(4,))
    f = chmy.Field(None, g, chmy.Center())
    with pytest.raises(chmy.ChmyError):
        f.desc()


def test_grid_metrics_and_coords_match_oracle():
    """UniformGrid spacing/inv_spacing and fma coordinates
    (uniform_axis.jl:7-19; test/test_grids.jl:10-175)."""
    g = chmy.UniformGrid(origin=(-1, -2), extent=(2, 4), dims=(5, 20))
    assert chmy.spacing(g) == (0.4, 0.2) and chmy.inv_spacing(g) == (2.5, 5.0)
    og = O.Grid((-1, -2), (2, 4), (5, 20))
    for loc in (chmy.Center(), chmy.Vertex()):
        for d in range(2):
            a = g.coords(loc, d, first=-1, count=g.dims[d] + 5)
            b = og.coords(loc.code, d, first=-1, count=g.dims[d] + 5)
            assert np.array_equal(a, b)
    assert math.isclose(chmy.coord(g, chmy.Center(), 0, 5), 0.8)


def test_field_layout():
    """Field storage n .+ 4H, x fastest, logical I at I .+ 2H (field.jl:6-18,56-62)."""
    g = chmy.UniformGrid(origin=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0), dims=(2, 2, 2))
    f = chmy.Field(None, g, (chmy.Center(), chmy.Vertex(), chmy.Center()), halo=(1, 0, 1))
    assert f.dims == (2, 3, 2) and f.pad == (6, 3, 6)
    assert tuple(f.data.shape) == (6, 3, 6)
    chmy.set_(f, lambda x, y, z: y + 0 * x, g)
    exp = np.array([0.0, 0.0, 0.5, 0.5, 1.0, 1.0, 0.0, 0.0, 0.5, 0.5, 1.0, 1.0])
    assert np.array_equal(f.interior().numpy().reshape(-1, order="F"), exp)
    v = chmy.VectorField(None, g)
    assert [tuple(l.code for l in c.loc) for c in v] == [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    t = chmy.TensorField(None, g)
    assert t._fields == ("xx", "yy", "zz", "xy", "xz", "yz")
    assert tuple(l.code for l in t.xz.loc) == (1, 0, 1)


def test_batch_regularisation():
    """batch(grid, f => bc...) per-dim/per-side FieldBatches with pruning of
    `nothing` and per-side tuples (batch.jl:72-157)."""
    g = chmy.UniformGrid(origin=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0), dims=(4, 4, 4))
    V = chmy.VectorField(None, g)
    D, N = chmy.Dirichlet, chmy.Neumann
    bs = chmy.batch(g, (V.x, {"x": D(), "y": N(), "z": N()}), (V.y, {"y": (D(1.0), D(-1.0))}),
                    (V.z, N()), exchange=(V.x, V.y, V.z))
    assert [len(bs[0][s].fields) for s in range(2)] == [2, 2]      # V.x, V.z in x
    assert [len(bs[1][s].fields) for s in range(2)] == [3, 3]
    assert bs[1][0].conditions[1].value == 1.0 and bs[1][1].conditions[1].value == -1.0
    assert isinstance(chmy.batch(g, (V.x, None))[0][0], chmy.EmptyBatch)
    with pytest.raises(ValueError):
        chmy.batch(g, (V.x, {"w": D()}))


def test_launcher_boxes():
    """worksize n .+ 2 at Offset(-1); inner/outer boxes (KernelLaunch.jl:40-87)."""
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(30, 20))
    L = chmy.Launcher(None, g, outer_width=(4, 3))
    assert L.worksize == (32, 22)
    assert L.inner_worksize() == (24, 16)
    assert L.outer_worksize(1) == (32, 3) and L.outer_offset(1, 2) == (0, 19)
    assert L.outer_worksize(0) == (4, 16) and L.outer_offset(0, 1) == (0, 3)
    # the slabs and the inner box tile the launcher box exactly
    cover = np.zeros(L.worksize, int)
    cover[4:28, 3:19] += 1
    for D in range(2):
        for S in (1, 2):
            o, w = L.outer_offset(D, S), L.outer_worksize(D)
            cover[o[0]:o[0] + w[0], o[1]:o[1] + w[1]] += 1
    assert np.all(cover == 1)


def test_error_messages_are_per_thread():
    """chmy_last_error is thread-local: threads that fail different calls at
    the same time each read their own message."""
    import threading
    L = _lib.lib()
    barrier = threading.Barrier(2)
    seen = {}

    def fail_stream():
        barrier.wait()
        for _ in range(200):
            assert L.chmy_stream_create(3, None) == -1
            seen.setdefault("stream", set()).add(L.chmy_last_error())

    def fail_grid():
        barrier.wait()
        for _ in range(200):
            assert L.chmy_diffusion_flux(None, None, None, None, 1.0, None) == -1
            seen.setdefault("grid", set()).add(L.chmy_last_error())

    ts = [threading.Thread(target=fail_stream), threading.Thread(target=fail_grid)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len(seen["stream"]) == 1 and b"stream is NULL" in next(iter(seen["stream"]))
    assert len(seen["grid"]) == 1 and b"grid is NULL" in next(iter(seen["grid"]))


def test_boundary_function_points_follow_the_reference():
    """batch() evaluates a BoundaryFunction at the points and locations
    first_order_boundary_condition.jl uses (host only, no device): a
    Vertex-located Dirichlet at the boundary nodes 1 and m with the field's
    location (:50-54), a Center-located Dirichlet at the ghosts 0 and m+1 with
    Center (:59-71), a Neumann at the ghosts with the flipped location
    (:76-84), over the tangential range 0..n+2."""
    from chmy import Center as Cc, Vertex as Vv
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 2.0), dims=(4, 5))
    calls = []

    def rec(grid, loc, dim, *I):
        calls.append((tuple(type(l).__name__[0] for l in loc), dim, I))
        return 0.0

    bf = chmy.BoundaryFunction(rec, discrete=True, reduce_dims=False)
    vx = chmy.Field(None, g, (Vv(), Cc()))
    chmy.batch(g, (vx, {"x": chmy.Dirichlet(bf), "y": chmy.Neumann(bf)}))
    x_calls = [c for c in calls if c[1] == 0]
    y_calls = [c for c in calls if c[1] == 1]
    assert {c[0] for c in x_calls} == {("V", "C")}
    assert sorted({c[2][0] for c in x_calls}) == [1, 5]               # boundary nodes 1, m = n+1
    assert sorted({c[2][1] for c in x_calls}) == list(range(0, 8))    # tangential 0..n_y+2
    assert {c[0] for c in y_calls} == {("C", "V")}                     # flipped
    assert sorted({c[2][1] for c in y_calls}) == [0, 6]               # ghosts 0, m+1
    assert sorted({c[2][0] for c in y_calls}) == list(range(0, 7))
    calls.clear()
    cf = chmy.Field(None, g, Cc())
    chmy.batch(g, (cf, {"y": chmy.Dirichlet(bf)}))
    assert {c[0] for c in calls} == {("C", "C")} and sorted({c[2][1] for c in calls}) == [0, 6]
    # reduced dims drop the boundary's own dim; a continuous function gets coordinates
    got = []
    red = chmy.BoundaryFunction(lambda y, s: got.append(y) or s * y, parameters=(2.0,))
    b = chmy.batch(g, (cf, {"x": chmy.Dirichlet(red)}))
    ys = g.coords(Cc(), 1, first=0, count=8)
    assert np.array_equal(np.array(got[:8]), ys)
    v = b[0][0].conditions[0].value
    assert np.array_equal(v.data.cpu().numpy()[1:9], 2.0 * ys)


def test_set_on_a_vector_field_sets_every_component():
    """set!(fs::NamedTuple, args...) (field.jl:144-145): each component is set
    with its own location's coordinates."""
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(4, 3))
    V = chmy.VectorField(None, g)
    chmy.set_(V, lambda x, y: x + 10.0 * y, g)
    for D, f in enumerate(V):
        xs = np.meshgrid(*[g.coords(f.loc, d) for d in range(2)], indexing="ij")
        assert np.array_equal(f.interior().numpy(), xs[0] + 10.0 * xs[1]), D
