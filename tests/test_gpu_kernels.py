"""GPU parity of every Launcher kernel against the CPU oracle (bitwise).

Inputs fill the whole padded storage with seeded random values so that reads
of halos and padding are exercised identically on both sides.  The bar is
bitwise equality in f64 and f32: both sides evaluate the same IEEE
expressions with no contraction and fma only at the reference's muladd
sites.
"""
import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc, Vertex as Vv
import oracle as O
from _helpers import NP2TORCH, assert_same, dev, ogrid, pair, vloc

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


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_partial_lerp_left_right(arch, T, nd):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(10 + nd)
    L = chmy.Launcher(arch, g)
    locs = [tuple(Cc() if (m >> d) & 1 == 0 else Vv() for d in range(nd)) for m in range(1 << nd)]
    for src_loc in locs:
        s, hs = pair(arch, g, og, src_loc, T, rng=rng)
        for D in range(nd):
            out_loc = tuple((Vv() if isinstance(l, Cc) else Cc()) if d == D else l for d, l in enumerate(src_loc))
            d_, hd = pair(arch, g, og, out_loc, T, fill=0)
            L(arch, g, (chmy.kernels.partial, (d_, s, D, g)))
            O.op("partial", og, hd, hs, dim=D)
            assert_same(d_, hd, f"∂ dim {D} src {src_loc}")
            for name, kern in (("left", chmy.kernels.left), ("right", chmy.kernels.right)):
                L(arch, g, (kern, (d_, s, D, g)))
                O.op(name, og, hd, hs, dim=D)
                assert_same(d_, hd, f"{name} dim {D}")
        for to in locs:
            d_, hd = pair(arch, g, og, to, T, fill=0)
            L(arch, g, (chmy.kernels.lerp, (d_, s, g)))
            O.op("lerp", og, hd, hs)
            assert_same(d_, hd, f"lerp {src_loc}->{to}")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [2, 3])
def test_divg_vmag(arch, T, nd):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(20 + nd)
    L = chmy.Launcher(arch, g)
    V = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    out, hout = pair(arch, g, og, Cc(), T, fill=0)
    L(arch, g, (chmy.kernels.divg, (out, [v for v, _ in V], g)))
    O.op("divg", og, hout, [h for _, h in V])
    assert_same(out, hout, "divg")
    L(arch, g, (chmy.kernels.vmag, (out, [v for v, _ in V], g)))
    O.op("vmag", og, hout, [h for _, h in V])
    assert_same(out, hout, "vmag")


def test_reference_operator_identities(arch):
    """test/test_grid_operators.jl:14-42,114-132 through the GPU Launcher:
    divg(V) == ∂xVx + ∂yVy + ∂zVz exactly; vmag of a constant 2-vector ≈ 3.4641."""
    g = chmy.UniformGrid(arch, origin=(-5.0, -5.0, -5.0), extent=(10.0, 10.0, 10.0), dims=(12, 10, 8))
    L = chmy.Launcher(arch, g)
    Ci = chmy.Field(arch, g, Cc())
    chmy.set_(Ci, lambda x, y, z: np.exp(-x ** 2 - y ** 2 - z ** 2), g)
    V = chmy.VectorField(arch, g)
    for D in range(3):
        L(arch, g, (chmy.kernels.partial, (V[D], Ci, D, g)))
    parts = [chmy.Field(arch, g, Cc()) for _ in range(3)]
    for D in range(3):
        L(arch, g, (chmy.kernels.partial, (parts[D], V[D], D, g)))
    C2 = chmy.Field(arch, g, Cc())
    L(arch, g, (chmy.kernels.divg, (C2, list(V), g)))
    C1 = (parts[0].interior() + parts[1].interior()) + parts[2].interior()
    assert bool((C2.interior() == C1).all())
    for D in range(3):
        chmy.set_(V[D], 2.0)
    L(arch, g, (chmy.kernels.vmag, (C2, list(V), g)))
    v = C2.interior().cpu().numpy()
    assert np.all(np.round(v, 4) == 3.4641)


# ---------------------------------------------------------------- boundary conditions
def _bc_specs(nd, value_fields):
    D, N = chmy.Dirichlet, chmy.Neumann
    names = "xyz"[:nd]
    specs = [D(), N(), D(2.0), N(-1.5)]
    mixed = {names[0]: (D(1.0), N(0.5))}
    if nd > 1:
        mixed[names[1]] = N()
    if nd > 2:
        mixed[names[2]] = (N(), D(-3.0))
    specs.append(mixed)
    if value_fields is not None:
        specs.append({names[0]: D(value_fields[0]), names[-1]: (N(value_fields[1]), D())})
    return specs


def _ospec(spec, nd, vmap):
    names = "xyz"[:nd]
    def conv(b):
        v = b.value
        if isinstance(v, chmy.Field):
            v = vmap[id(v)]
        return O.BC(b.kind, v)
    if isinstance(spec, chmy.FirstOrderBC):
        return conv(spec)
    out = {}
    for k, b in spec.items():
        d = names.index(k)
        out[d] = tuple(conv(x) for x in b) if isinstance(b, tuple) else conv(b)
    return out


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("nfields", [4, 9])
def test_bc_batchset_equals_per_side_launches(arch, T, nd, nfields):
    """chmy_bc_batchset (one launch for the whole set when at most 8 distinct
    fields, else one per dim) equals the reference's sequence of bc! calls,
    dims N..1, side 1 then side 2 (batch.jl:20-29), bitwise, for a batch that
    mixes locations and conditions per field."""
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(70 + nd + nfields)
    names = "xyz"[:nd]
    D, N = chmy.Dirichlet, chmy.Neumann
    conds = [D(), N(), D(1.25), N(-0.75)]
    pairs_a, pairs_b = [], []
    for q in range(nfields):
        loc = tuple(Cc() if (q >> d) & 1 == 0 else Vv() for d in range(nd))
        spec = {}
        for d, k in enumerate(names):
            c1, c2 = conds[(q + d) % 4], conds[(q + 2 * d + 1) % 4]
            spec[k] = (c1, c2) if (q + d) % 3 else c1
        fa, _ = pair(arch, g, og, loc, T, rng=rng)
        fb = chmy.Field(arch, g, loc, fa.data.dtype)
        fb.data.copy_(fa.data)
        pairs_a.append((fa, spec))
        pairs_b.append((fb, spec))
    ba = chmy.batch(g, *pairs_a)
    bb = chmy.batch(g, *pairs_b)
    ba.apply()
    for d in reversed(range(nd)):
        for s in (1, 2):
            bb.apply_side(d, s)
    bits = np.int64 if T == np.float64 else np.int32
    for q, ((fa, spec), (fb, _)) in enumerate(zip(pairs_a, pairs_b)):
        assert np.array_equal(dev(fa).view(bits), dev(fb).view(bits)), f"field {q} spec {spec}"


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_bc_batches(arch, T, nd):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(30 + nd)
    vmap = {}
    vfields = None
    if nd >= 2:
        # value fields on the (nd-1)-D tangential grid (Center sizes, halo 1)
        tg = chmy.UniformGrid(origin=g.origin[1:], extent=g.extent[1:], dims=g.dims[1:])
        tg0 = chmy.UniformGrid(origin=g.origin[:-1], extent=g.extent[:-1], dims=g.dims[:-1])
        otg, otg0 = ogrid(tg), ogrid(tg0)
        a, ha = pair(arch, tg, otg, Cc(), T, rng=rng)
        b, hb = pair(arch, tg0, otg0, Cc(), T, rng=rng)
        vfields = (a, b)
        vmap = {id(a): ha, id(b): hb}
    locs = [tuple(Cc() if (m >> d) & 1 == 0 else Vv() for d in range(nd)) for m in range(1 << nd)]
    for spec in _bc_specs(nd, vfields):
        fs = [pair(arch, g, og, l, T, rng=rng) for l in locs]
        chmy.bc_(arch, g, *[(f, spec) for f, _ in fs])
        O.bc_apply(og, [(h, _ospec(spec, nd, vmap)) for _, h in fs])
        for (f, h), l in zip(fs, locs):
            assert_same(f, h, f"bc {spec} loc {l}")


def test_reference_bc_ghost_relations_3d(arch):
    """test/test_boundary_conditions.jl:140-205 on the GPU: 8x8x6, loc (C,V,C)."""
    import math
    g = chmy.UniformGrid(arch, origin=(-math.pi,) * 3, extent=(2 * math.pi,) * 3, dims=(8, 8, 6))
    f = chmy.Field(arch, g, (Cc(), Vv(), Cc()))
    s = slice(1, -1)
    chmy.set_(f, 1.0)
    chmy.bc_(arch, g, (f, chmy.Dirichlet(2.0)))
    fi = f.interior(with_halo=True).cpu().numpy()
    assert np.allclose(fi[0, s, s], -fi[1, s, s] + 4) and np.allclose(fi[-1, s, s], -fi[-2, s, s] + 4)
    assert np.all(fi[s, 1, s] == 2) and np.all(fi[s, -2, s] == 2)
    assert np.allclose(fi[s, s, 0], -fi[s, s, 1] + 4) and np.allclose(fi[s, s, -1], -fi[s, s, -2] + 4)
    chmy.set_(f, 1.0)
    chmy.bc_(arch, g, (f, chmy.Neumann(2.0)))
    fi = f.interior(with_halo=True).cpu().numpy()
    dx, dy, dz = chmy.spacing(g)
    assert np.allclose((fi[1, s, s] - fi[0, s, s]) / dx, 2.0)
    assert np.allclose((fi[s, -1, s] - fi[s, -2, s]) / dy, 2.0)
    assert np.allclose((fi[s, s, 1] - fi[s, s, 0]) / dz, 2.0)


# ---------------------------------------------------------------- diffusion (config: 256² x 1000 steps)
def test_diffusion_2d_256_1000_steps(arch):
    """examples/diffusion_2d.jl loop at 256² on [-1,1]², C0 = exp(-(x²+y²)/0.1) +
    1e-3·noise, χ = 1, Δt = min(Δ)²/4.1, Neumann(0), 1000 steps: every stored
    value (interior + post-BC halos) bitwise equal to the oracle."""
    n, steps = 256, 1000
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(n, n))
    og = ogrid(g)
    L = chmy.Launcher(arch, g, sync=False)
    rng = np.random.default_rng(1234)
    xs = og.coords(O.CENTER, 0)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    C0 = np.exp(-(X ** 2 + Y ** 2) / 0.1) + 1e-3 * rng.random((n, n))
    Cf = chmy.Field(arch, g, Cc())
    q = chmy.VectorField(arch, g)
    hC = O.HField(og, O.CENTER)
    hq = [O.HField(og, (O.VERTEX, O.CENTER)), O.HField(og, (O.CENTER, O.VERTEX))]
    chmy.set_(Cf, C0)
    hC.set_interior(C0)
    bcs = chmy.batch(g, (Cf, chmy.Neumann()), exchange=Cf)
    chmy.bc_(arch, g, bcs)
    O.bc_apply(og, [(hC, O.Neumann())])
    dt = min(chmy.spacing(g)) ** 2 / 1.0 / 4.1
    for _ in range(steps):
        L(arch, g, (chmy.kernels.compute_q, (q, Cf, 1.0, g)))
        L(arch, g, (chmy.kernels.update_C, (Cf, q, dt, g)), bc=bcs)
        O.compute_q(og, hq, hC, 1.0)
        O.update_C(og, hC, hq, dt)
        O.bc_apply(og, [(hC, O.Neumann())])
    chmy.synchronize()
    assert_same(Cf, hC, "C after 1000 steps")
    assert_same(q.x, hq[0], "qx")


def test_stream_and_zero_fill(arch):
    """chmy_stream_create/destroy and chmy_memset_zero (unaligned head and
    tail, 16-byte body) on a device buffer."""
    import ctypes as C
    import torch
    from chmy import _lib
    s = C.c_void_p()
    _lib.call("chmy_stream_create", 0, C.byref(s))
    buf = torch.full((4099,), 0x5A, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    _lib.call("chmy_memset_zero", C.c_void_p(buf.data_ptr() + 3), 4000, s)
    _lib.call("chmy_stream_sync", s)
    h = buf.cpu().numpy()
    assert (h[:3] == 0x5A).all() and (h[3:4003] == 0).all() and (h[4003:] == 0x5A).all()
    _lib.call("chmy_stream_destroy", s)


def test_diffusion_iterate_256_1000_steps(arch):
    """The same 256² x 1000-step diffusion loop as one native call (chmy_diffusion_iterate:
    fused step into the other buffer + Neumann batch, 1000 times): every
    stored value bitwise equal to the oracle's loop."""
    import ctypes as C
    from chmy import _lib
    from chmy.kernels import _arr
    n, steps = 256, 1000
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(n, n))
    og = ogrid(g)
    rng = np.random.default_rng(1234)
    xs = og.coords(O.CENTER, 0)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    C0 = np.exp(-(X ** 2 + Y ** 2) / 0.1) + 1e-3 * rng.random((n, n))
    Cs = [chmy.Field(arch, g, Cc()), chmy.Field(arch, g, Cc())]
    q = chmy.VectorField(arch, g)
    hC = O.HField(og, O.CENTER)
    hq = [O.HField(og, (O.VERTEX, O.CENTER)), O.HField(og, (O.CENTER, O.VERTEX))]
    chmy.set_(Cs[0], C0)
    chmy.bc_(arch, g, (Cs[0], chmy.Neumann()))
    Cs[1].data.copy_(Cs[0].data)
    hC.set_interior(C0)
    O.bc_apply(og, [(hC, O.Neumann())])
    dt = min(chmy.spacing(g)) ** 2 / 1.0 / 4.1
    counts = (C.c_int32 * 6)(1, 1, 1, 1, 0, 0)
    bcs = (_lib.ChmyBC * 4)(*[_lib.ChmyBC(_lib.NEUMANN, _lib.VALUE_ZERO, 0.0, None) for _ in range(4)])
    par = C.c_int(0)
    _lib.call("chmy_diffusion_iterate", C.byref(g.desc()), C.byref(Cs[0].desc()), C.byref(Cs[1].desc()), _arr(q),
              counts, bcs, 1.0, dt, steps, C.byref(par), chmy.current_stream())
    for _ in range(steps):
        O.compute_q(og, hq, hC, 1.0)
        O.update_C(og, hC, hq, dt)
        O.bc_apply(og, [(hC, O.Neumann())])
    chmy.synchronize()
    assert par.value == 0
    assert_same(Cs[par.value], hC, "C after 1000 steps")
    assert_same(q.x, hq[0], "qx")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("mix", ["neumann0", "mixed"])
def test_diffusion_iterate_folded_batch_equals_step_and_batchset(arch, T, nd, mix):
    """chmy_diffusion_iterate with one scalar rule on every (dim, side) runs
    the batch inside the step's launch; 5 steps of it equal chmy_diffusion_step
    + chmy_bc_batchset per step bitwise (both buffers and q), on random storage
    (signed zeros included), for zero-flux Neumann and for a mix of Dirichlet
    and Neumann with non-zero values."""
    import ctypes as C
    from chmy import _lib
    from chmy.kernels import _arr
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(60 + nd)
    if mix == "neumann0":
        rules = [(_lib.NEUMANN, _lib.VALUE_ZERO, 0.0)] * (2 * nd)
    else:
        pool = [(_lib.DIRICHLET, _lib.VALUE_SCALAR, 0.7), (_lib.NEUMANN, _lib.VALUE_SCALAR, -1.3),
                (_lib.DIRICHLET, _lib.VALUE_ZERO, 0.0), (_lib.NEUMANN, _lib.VALUE_SCALAR, 0.25),
                (_lib.NEUMANN, _lib.VALUE_ZERO, 0.0), (_lib.DIRICHLET, _lib.VALUE_SCALAR, -2.0)]
        rules = pool[:2 * nd]
    counts = (C.c_int32 * 6)(*([1] * (2 * nd) + [0] * (6 - 2 * nd)))
    bcs = (_lib.ChmyBC * (2 * nd))(*[_lib.ChmyBC(k, vk, v, None) for k, vk, v in rules])
    sets = []
    for _ in range(2):
        Cs = [chmy.Field(arch, g, Cc(), NP2TORCH[T]), chmy.Field(arch, g, Cc(), NP2TORCH[T])]
        q = [chmy.Field(arch, g, vloc(D, nd), NP2TORCH[T]) for D in range(nd)]
        sets.append((Cs, q))
    for f in sets[0][0] + sets[0][1]:
        a = rng.uniform(-1.0, 1.0, tuple(f.data.shape))
        a[a < -0.95] = -0.0   # signed zeros meet the batch's fma steps
        f.data.copy_(torch.from_numpy(a).to(f.data.dtype))
    for a, b in zip(sets[1][0] + sets[1][1], sets[0][0] + sets[0][1]):
        a.data.copy_(b.data)
    dt = 0.2 * min(chmy.spacing(g)) ** 2
    steps = 5
    par = C.c_int(0)
    Cs, q = sets[0]
    _lib.call("chmy_diffusion_iterate", C.byref(g.desc()), C.byref(Cs[0].desc()), C.byref(Cs[1].desc()), _arr(q),
              counts, bcs, 0.9, dt, steps, C.byref(par), chmy.current_stream())
    Cr, qr = sets[1]
    box = chmy.kernels._box((0,) * nd, tuple(n + 2 for n in g.dims))
    p = 0
    for _ in range(steps):
        _lib.call("chmy_diffusion_step", C.byref(g.desc()), C.byref(box), C.byref(Cr[p].desc()),
                  C.byref(Cr[1 - p].desc()), _arr(qr), 0.9, dt, chmy.current_stream())
        _lib.call("chmy_bc_batchset", C.byref(g.desc()), counts, _arr([Cr[1 - p]] * (2 * nd)), bcs,
                  chmy.current_stream())
        p = 1 - p
    chmy.synchronize()
    assert par.value == p
    bits = np.int64 if T == np.float64 else np.int32
    for k, (a, b) in enumerate(zip(Cs + q, Cr + qr)):
        assert np.array_equal(dev(a).view(bits), dev(b).view(bits)), f"array {k}"


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_diffusion_step_equals_two_kernels(arch, T, nd):
    """chmy_diffusion_step (compute_q! + update_C! in one launch into a second
    buffer) gives the stored q and the new C of the two launches, bitwise, on
    random storage over the Launcher box and a sub-box."""
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(40 + nd)
    Cf, hC = pair(arch, g, og, (Cc(),) * nd, T, rng=rng)
    qs = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    dt = 0.37 * min(chmy.spacing(g)) ** 2
    for box in (None, ((1,) * nd, tuple(n - 1 for n in g.dims))):
        Co = chmy.Field(arch, g, Cc(), Cf.data.dtype)
        q = [chmy.Field(arch, g, f.loc, f.data.dtype) for f, _ in qs]
        for a, (b, _) in zip(q, qs):
            a.data.copy_(b.data)
        lo, ws = ((0,) * nd, tuple(n + 2 for n in g.dims)) if box is None else box
        chmy.kernels.diffusion_step.launch(g, lo, ws, Cf, Co, q, 1.3, dt, g)
        # reference: the two launches on copies
        C2 = chmy.Field(arch, g, Cc(), Cf.data.dtype)
        C2.data.copy_(Cf.data)
        q2 = [chmy.Field(arch, g, f.loc, f.data.dtype) for f, _ in qs]
        for a, (b, _) in zip(q2, qs):
            a.data.copy_(b.data)
        chmy.kernels.compute_q.launch(g, lo, ws, q2, C2, 1.3, g)
        chmy.kernels.update_C.launch(g, lo, ws, C2, q2, dt, g)
        chmy.synchronize()
        # C_out holds the new values on the box and nothing elsewhere (zeros)
        got, want = dev(Co), dev(C2)
        sl = tuple(slice(1 + l, 1 + l + w) for l, w in zip(reversed(lo), reversed(ws)))
        bits = np.int64 if T == np.float64 else np.int32
        assert np.array_equal(got[sl].view(bits), want[sl].view(bits)), f"box {box}"
        for a, b in zip(q, q2):
            assert np.array_equal(dev(a).view(bits), dev(b).view(bits)), f"q box {box}"


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_thermal_step_equals_two_kernels(arch, T, nd):
    """chmy_thermal_step (update_thermal_flux! + update_thermal! in one launch
    into a second T buffer) gives the stored qT and the new T of the two
    launches, bitwise, on random storage with signed V (both upwind branches)
    over the Launcher box and a sub-box."""
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(50 + nd)
    Tf, _ = pair(arch, g, og, (Cc(),) * nd, T, rng=rng)
    To, _ = pair(arch, g, og, (Cc(),) * nd, T, rng=rng)
    V = [pair(arch, g, og, vloc(D, nd), T, rng=rng)[0] for D in range(nd)]
    qs = [pair(arch, g, og, vloc(D, nd), T, rng=rng)[0] for D in range(nd)]
    lam, dt = 0.7, 0.21 * min(chmy.spacing(g)) ** 2
    bits = np.int64 if T == np.float64 else np.int32
    for box in (None, ((1,) * nd, tuple(n - 1 for n in g.dims))):
        Tn = chmy.Field(arch, g, Cc(), Tf.data.dtype)
        q = [chmy.Field(arch, g, f.loc, f.data.dtype) for f in qs]
        for a, b in zip(q, qs):
            a.data.copy_(b.data)
        lo, ws = ((0,) * nd, tuple(n + 2 for n in g.dims)) if box is None else box
        chmy.kernels.thermal_step.launch(g, lo, ws, Tf, Tn, To, q, V, lam, dt, g)
        T2 = chmy.Field(arch, g, Cc(), Tf.data.dtype)
        T2.data.copy_(Tf.data)
        q2 = [chmy.Field(arch, g, f.loc, f.data.dtype) for f in qs]
        for a, b in zip(q2, qs):
            a.data.copy_(b.data)
        chmy.kernels.update_thermal_flux.launch(g, lo, ws, q2, T2, V, lam, g)
        chmy.kernels.update_thermal.launch(g, lo, ws, T2, To, q2, dt, g)
        chmy.synchronize()
        sl = tuple(slice(1 + l, 1 + l + w) for l, w in zip(reversed(lo), reversed(ws)))
        assert np.array_equal(dev(Tn)[sl].view(bits), dev(T2)[sl].view(bits)), f"box {box}"
        for a, b in zip(q, q2):
            assert np.array_equal(dev(a).view(bits), dev(b).view(bits)), f"qT box {box}"
    with pytest.raises(chmy.ChmyError):
        chmy.kernels.thermal_step.launch(g, lo, ws, Tf, Tf, To, q, V, lam, dt, g)


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("dims", [(130, 19, 45), (64, 6, 30), (65, 9, 2), (3, 2, 70)])
def test_thermal_step_3d_march_equals_two_kernels(arch, T, dims):
    """The 3D thermal step marches z over chunks of up to 16 planes in 64 x 4
    blocks: several x tiles, y tiles and z chunks, exactly full and one-over
    tiles, boxes that start inside the grid and end before its last index
    (the right faces' fluxes then come from qT's storage), against the two
    launches, bitwise, T and every qT point."""
    g = chmy.UniformGrid(origin=(-0.5, -0.5, -0.5), extent=(1.0, 0.9, 1.2), dims=dims)
    og = ogrid(g)
    rng = np.random.default_rng(sum(dims))
    Tf, _ = pair(arch, g, og, (Cc(),) * 3, T, rng=rng)
    To, _ = pair(arch, g, og, (Cc(),) * 3, T, rng=rng)
    V = [pair(arch, g, og, vloc(D, 3), T, rng=rng)[0] for D in range(3)]
    qs = [pair(arch, g, og, vloc(D, 3), T, rng=rng)[0] for D in range(3)]
    lam, dt = 0.7, 0.21 * min(chmy.spacing(g)) ** 2
    bits = np.int64 if T == np.float64 else np.int32
    n = g.dims
    boxes = [((0, 0, 0), tuple(m + 2 for m in n)), ((1, 1, 1), tuple(m for m in n)),
             ((2, 0, 1), (n[0] - 1, n[1] + 1, n[2]))]
    for lo, ws in boxes:
        Tn = chmy.Field(arch, g, Cc(), Tf.data.dtype)
        q = [chmy.Field(arch, g, f.loc, f.data.dtype) for f in qs]
        for a, b in zip(q, qs):
            a.data.copy_(b.data)
        chmy.kernels.thermal_step.launch(g, lo, ws, Tf, Tn, To, q, V, lam, dt, g)
        T2 = chmy.Field(arch, g, Cc(), Tf.data.dtype)
        T2.data.copy_(Tf.data)
        q2 = [chmy.Field(arch, g, f.loc, f.data.dtype) for f in qs]
        for a, b in zip(q2, qs):
            a.data.copy_(b.data)
        chmy.kernels.update_thermal_flux.launch(g, lo, ws, q2, T2, V, lam, g)
        chmy.kernels.update_thermal.launch(g, lo, ws, T2, To, q2, dt, g)
        chmy.synchronize()
        sl = tuple(slice(1 + l, 1 + l + w) for l, w in zip(reversed(lo), reversed(ws)))
        assert np.array_equal(dev(Tn)[sl].view(bits), dev(T2)[sl].view(bits)), f"box {lo} {ws}"
        assert not np.any(dev(Tn)[sl] == 0)
        for a, b in zip(q, q2):
            assert np.array_equal(dev(a).view(bits), dev(b).view(bits)), f"qT box {lo} {ws}"


def test_empty_box_is_a_no_op(arch):
    """A box with no points along some dim launches nothing and returns 0, as
    a KernelAbstractions launch over an empty worksize does: every field keeps
    its bits."""
    g = grids(3)
    og = ogrid(g)
    rng = np.random.default_rng(77)
    Tf, _ = pair(arch, g, og, (Cc(),) * 3, np.float64, rng=rng)
    Tn, _ = pair(arch, g, og, (Cc(),) * 3, np.float64, rng=rng)
    To, _ = pair(arch, g, og, (Cc(),) * 3, np.float64, rng=rng)
    V = [pair(arch, g, og, vloc(D, 3), np.float64, rng=rng)[0] for D in range(3)]
    q = [pair(arch, g, og, vloc(D, 3), np.float64, rng=rng)[0] for D in range(3)]
    fields = [Tf, Tn, To] + V + q
    before = [dev(f).copy() for f in fields]
    for ws in ((0, 5, 5), (5, 0, 5), (5, 5, 0)):
        chmy.kernels.thermal_step.launch(g, (1, 1, 1), ws, Tf, Tn, To, q, V, 0.7, 1e-3, g)
        chmy.kernels.update_thermal_flux.launch(g, (1, 1, 1), ws, q, Tf, V, 0.7, g)
        chmy.kernels.update_thermal.launch(g, (1, 1, 1), ws, Tf, To, q, 1e-3, g)
    chmy.synchronize()
    for a, f in zip(before, fields):
        assert np.array_equal(a.view(np.int64), dev(f).view(np.int64))


@pytest.mark.parametrize("ow", [(16, 8), (3, 5)])
def test_launcher_outer_width_matches(arch, ow):
    """Launcher(outer_width): inner box + 2N slabs with per-side bc
    (KernelLaunch.jl:152-183) give the same fields as the plain launch."""
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(62, 40))
    rng = np.random.default_rng(5)
    res = []
    for width in (None, ow):
        L = chmy.Launcher(arch, g, outer_width=width)
        Cf = chmy.Field(arch, g, Cc())
        q = chmy.VectorField(arch, g)
        chmy.set_(Cf, rng.random((62, 40)) if not res else res[0][1])
        c0 = Cf.interior().cpu().numpy().copy()
        bcs = chmy.batch(g, (Cf, chmy.Neumann()))
        chmy.bc_(arch, g, bcs)
        for _ in range(20):
            L(arch, g, (chmy.kernels.compute_q, (q, Cf, 1.0, g)))
            L(arch, g, (chmy.kernels.update_C, (Cf, q, 1e-4, g)), bc=bcs)
        res.append((dev(Cf), c0))
    assert np.array_equal(res[0][0], res[1][0])


# ---------------------------------------------------------------- Stokes kernels
def _stokes_pairs(arch, g, og, T, rng):
    nd = g.nd
    tl = list(chmy.tensor_locations(nd).values())
    F = dict(
        Pr=pair(arch, g, og, Cc(), T, rng=rng), dV=pair(arch, g, og, Cc(), T, rng=rng),
        V=[pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)],
        rV=[pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)],
        tau=[pair(arch, g, og, l, T, rng=rng) for l in tl],
        tauo=[pair(arch, g, og, l, T, rng=rng) for l in tl],
        eta=pair(arch, g, og, Cc(), T, rng=rng, lo=1e-3, hi=3.0),
        rg=pair(arch, g, og, vloc(nd - 1, nd), T, rng=rng),
    )
    return F


def _d(F, k):
    v = F[k]
    return [a for a, _ in v] if isinstance(v, list) else v[0]


def _h(F, k):
    v = F[k]
    return [b for _, b in v] if isinstance(v, list) else v[1]


def _check_all(F, what):
    for k, v in F.items():
        for i, (a, b) in enumerate(v if isinstance(v, list) else [v]):
            assert_same(a, b, f"{what}: {k}[{i}]")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("visc", ["scalar", "field"])
def test_stress_velocity(arch, T, nd, visc):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(40 + nd)
    F = _stokes_pairs(arch, g, og, T, rng)
    eml, heml = pair(arch, g, og, Cc(), T, fill=0)
    L = chmy.Launcher(arch, g)
    G, dt, eta = 1.0, 0.37, 0.8
    eta_ve = 1.0 / (1.0 / eta + 1.0 / (G * dt))
    p = O.pt_params(og)
    if visc == "field":
        L(arch, g, (chmy.kernels.maxloc, (eml, _d(F, "eta"))))
        O.maxloc(og, heml, _h(F, "eta"))
        assert_same(eml, heml, "maxloc")
        # the Launcher box touches eta_ml at I-e_D: fill the rest of its storage identically
        heml.arr[...] = dev(eml)
    ed = _d(F, "eta") if visc == "field" else eta
    eh = _h(F, "eta") if visc == "field" else eta
    L(arch, g, (chmy.kernels.update_stress, (_d(F, "tau"), _d(F, "Pr"), _d(F, "dV"), _d(F, "V"), _d(F, "tauo"),
                                             ed, eta_ve, G, dt, p["dtau_Pr"], p["dtau_r"], g)))
    O.update_stress(og, _h(F, "tau"), _h(F, "Pr"), _h(F, "dV"), _h(F, "V"), _h(F, "tauo"), eh, eta_ve, G, dt,
                    p["dtau_Pr"], p["dtau_r"])
    _check_all(F, "after update_stress")
    if visc == "field":
        L(arch, g, (chmy.kernels.update_velocity_visc, (_d(F, "V"), _d(F, "rV"), _d(F, "Pr"), _d(F, "tau"),
                                                        _d(F, "rg"), eml, G, dt, p["nudtau"], g)))
        O.update_velocity(og, _h(F, "V"), _h(F, "rV"), _h(F, "Pr"), _h(F, "tau"), _h(F, "rg"), 0.0, p["nudtau"],
                          eta_ml=heml, G=G, dt=dt)
    else:
        L(arch, g, (chmy.kernels.update_velocity, (_d(F, "V"), _d(F, "rV"), _d(F, "Pr"), _d(F, "tau"),
                                                   _d(F, "rg"), eta_ve, p["nudtau"], g)))
        O.update_velocity(og, _h(F, "V"), _h(F, "rV"), _h(F, "Pr"), _h(F, "tau"), _h(F, "rg"), eta_ve, p["nudtau"])
    _check_all(F, "after update_velocity")


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd", [2, 3])
def test_update_old_and_thermal(arch, T, nd):
    g = grids(nd)
    og = ogrid(g)
    rng = np.random.default_rng(50 + nd)
    L = chmy.Launcher(arch, g)
    tl = list(chmy.tensor_locations(nd).values())
    tau = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    tauo = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    Tf, hT = pair(arch, g, og, Cc(), T, rng=rng)
    To, hTo = pair(arch, g, og, Cc(), T, rng=rng)
    L(arch, g, (chmy.kernels.update_old, (Tf, [a for a, _ in tau], To, [a for a, _ in tauo])))
    O.update_old(og, hT, [b for _, b in tau], hTo, [b for _, b in tauo])
    for (a, b) in tauo + [(To, hTo)]:
        assert_same(a, b, "update_old")
    V = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    qT = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    L(arch, g, (chmy.kernels.update_thermal_flux, ([a for a, _ in qT], Tf, [a for a, _ in V], 0.37, g)))
    O.thermal_flux(og, [b for _, b in qT], hT, [b for _, b in V], 0.37)
    for a, b in qT:
        assert_same(a, b, "thermal flux")
    L(arch, g, (chmy.kernels.update_thermal, (Tf, To, [a for a, _ in qT], 1e-3, g)))
    O.thermal(og, hT, hTo, [b for _, b in qT], 1e-3)
    assert_same(Tf, hT, "thermal update")


# ---------------------------------------------------------------- reductions
@pytest.mark.parametrize("T", TYPES)
def test_reductions(arch, T):
    g = grids(3)
    og = ogrid(g)
    rng = np.random.default_rng(60)
    f, h = pair(arch, g, og, (Vv(), Cc(), Vv()), T, rng=rng, lo=-5, hi=5)
    assert chmy.kernels.maxabs(f) == O.maxabs(h)
    s = chmy.kernels.sumsq(f)
    assert np.isclose(s, O.sumsq(h), rtol=1e-12 if T == np.float64 else 1e-6)
    assert s == chmy.kernels.sumsq(f)  # deterministic order
    # NaN propagates
    f.interior()[3, 2, 1] = float("nan")
    assert np.isnan(chmy.kernels.maxabs(f))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("dims", [(600, 3, 2), (257,), (1000, 7), (5, 5, 5), (256, 9), (513, 1, 3)])
def test_maxabs_items(arch, dtype, dims):
    """max|x| and the sum of squares over the (row, 256-element segment)
    items the kernels split the box into: rows of one to four segments, item counts that are not a
    multiple of the items a wave takes, and the maximum planted at the first
    and last interior element and in a later segment of a middle row; halo
    values above the maximum are ignored; exact against numpy."""
    nd = len(dims)
    g = chmy.UniformGrid(arch, origin=(0.0,) * nd, extent=(1.0,) * nd, dims=dims)
    f = chmy.Field(arch, g, (Cc(),) * nd, dtype)
    rng = np.random.default_rng(sum(dims))
    n = dims   # interior() indexes (x, y, z)
    spots = [(0,) * nd, tuple(m - 1 for m in n), (min(n[0] - 1, 300),) + tuple(m // 2 for m in n[1:])]
    for q, spot in enumerate(spots):
        base = torch.from_numpy(rng.uniform(-1.0, 1.0, tuple(f.data.shape))).to(dtype)
        base.fill_(100.0)                                # halo
        inner = torch.from_numpy(rng.uniform(-1.0, 1.0, n)).to(dtype)
        inner[spot] = -7.5 if q % 2 else 7.5
        f.data.copy_(base)
        f.interior().copy_(inner.to(f.data.device))
        want = float(np.max(np.abs(inner.double().numpy())))
        assert want == 7.5
        assert chmy.kernels.maxabs(f) == want, (dims, spot)
        ss = chmy.kernels.sumsq(f)   # same items; a fixed order, so repeatable bit for bit
        assert np.isclose(ss, float(np.sum(inner.double().numpy() ** 2)), rtol=1e-12, atol=0), (dims, spot)
        assert ss == chmy.kernels.sumsq(f)
    f.interior()[spots[-1]] = float("nan")
    assert np.isnan(chmy.kernels.maxabs(f))
