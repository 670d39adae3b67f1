"""Two PT iterations per pass (2D and 3D, the free-slip V batch applied in-kernel
between them) against two single fused iterations with chmy_bc_batchset
between and after them, bitwise, on seeded random storage.

chmy_stokes_iterate takes iterations in pairs when the V batch is the
free-slip batch (each component Dirichlet on its own dim, Neumann on the
others; scalar values) and no residual fields are requested (in 3D not with
rho g and forcing together); the reference here is the single-iteration path,
driven one iteration at a time (chmy_stokes_iter + chmy_bc_batchset), which
tests/test_gpu_fused.py pins to the oracle.  Grids cover partial x tiles and
y chunks, chunks of one row, and the smallest grid the pass takes (2x2).
chmy_stokes_iter2 (one pass, no batch after it) is checked the same way, and
chmy_stokes_iter2_first's V of the first iteration against one iteration.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy import _lib
from chmy.kernels import _arr
from chmy.stokes import pt_params

pytestmark = pytest.mark.gpu

ITER_EXACT, ITER_PAIRS, ITER_SINGLE = 2, 4, 8


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _batch3(v):
    """(counts, comp, conds) for the 3D free-slip batch, v[component][dim][side]."""
    D, N, S = _lib.DIRICHLET, _lib.NEUMANN, _lib.VALUE_SCALAR
    rows = [(c, D if c == d else N, v[c][d][s]) for d in range(3) for s in range(2) for c in range(3)]
    counts = (C.c_int32 * 6)(*([3] * 6))
    comp = (C.c_int32 * 18)(*[c for c, _, _ in rows])
    conds = (_lib.ChmyBC * 18)(*[_lib.ChmyBC(k, S, float(x), None) for _, k, x in rows])
    return counts, comp, conds


def _batch(vals):
    """(counts, comp, conds) in (dim, side) order for the 2D free-slip batch."""
    (ax1, ax2), (qy1, qy2), (qx1, qx2), (by1, by2) = vals
    D, N, S = _lib.DIRICHLET, _lib.NEUMANN, _lib.VALUE_SCALAR
    rows = [(0, D, ax1), (1, N, qx1),      # dim x, side 1
            (0, D, ax2), (1, N, qx2),      # dim x, side 2
            (0, N, qy1), (1, D, by1),      # dim y, side 1
            (0, N, qy2), (1, D, by2)]      # dim y, side 2
    counts = (C.c_int32 * 6)(2, 2, 2, 2, 0, 0)
    comp = (C.c_int32 * 8)(*[c for c, _, _ in rows])
    conds = (_lib.ChmyBC * 8)(*[_lib.ChmyBC(k, S, float(v), None) for _, k, v in rows])
    return counts, comp, conds


class _Set:
    def __init__(self, arch, g, dt):
        self.V = chmy.VectorField(arch, g, dt)
        self.tau = chmy.TensorField(arch, g, dt)
        self.Pr = chmy.Field(arch, g, Cc(), dt)

    def fields(self):
        return list(self.V) + list(self.tau) + [self.Pr]


def _run(arch, g, T, exact, visc, forcing, vals, niter, pairs, seed, rhog=False, mode=None):
    dt = torch.float64 if T == np.float64 else torch.float32
    rng = np.random.default_rng(seed)
    fill = lambda f, lo=-1.0, hi=1.0: f.data.copy_(
        torch.from_numpy(rng.uniform(lo, hi, tuple(f.data.shape))).to(f.data.dtype))
    sets = [_Set(arch, g, dt), _Set(arch, g, dt)]
    for f in sets[0].fields():
        fill(f)
    for a, b in zip(sets[1].fields(), sets[0].fields()):
        a.data.copy_(b.data)
    fdt = torch.bfloat16 if forcing == "bf16" else dt
    told = chmy.TensorField(arch, g, fdt)
    for f in told:
        fill(f)
    eta = eml = None
    if visc == "field":
        eta, eml = chmy.Field(arch, g, Cc(), dt), chmy.Field(arch, g, Cc(), dt)
        fill(eta, 1e-3, 3.0)
        fill(eml, 1e-3, 3.0)
    fV = fP = None
    if forcing:
        fV, fP = chmy.VectorField(arch, g, fdt), chmy.Field(arch, g, Cc(), fdt)
        for f in list(fV) + [fP]:
            fill(f, -0.5, 0.5)
    rg = None
    if rhog:   # rho g at the last component's location (update_velocity!'s -rho g)
        from chmy.fields import vector_location
        rg = chmy.Field(arch, g, vector_location(g.nd - 1, g.nd), dt)
        fill(rg, -0.5, 0.5)
    divV = chmy.Field(arch, g, Cc(), dt)
    rV = chmy.VectorField(arch, g, dt)
    counts, comp, conds = _batch(vals) if g.nd == 2 else _batch3(vals)
    G, dtp, eta_s = 1.0, 0.37, 0.8
    eve = 1.0 / (1.0 / eta_s + 1.0 / (G * dtp))
    p = pt_params(g)
    flags = ITER_EXACT if exact else 0
    par = C.c_int(0)
    args = lambda a, b: (C.byref(g.desc()), _arr(list(a.V)), _arr(list(a.tau)), C.byref(a.Pr.desc()),
                         _arr(list(b.V)), _arr(list(b.tau)), C.byref(b.Pr.desc()), _arr(list(told)),
                         C.byref(eta.desc()) if eta is not None else None,
                         C.byref(eml.desc()) if eml is not None else None,
                         C.byref(rg.desc()) if rg is not None else None,
                         _arr(list(fV)) if fV is not None else None,
                         C.byref(fP.desc()) if fP is not None else None, C.byref(divV.desc()), _arr(list(rV)))
    scal = (eta_s, eve, G, dtp, p["dtau_Pr"], p["dtau_r"], p["nudtau"])
    if mode == "iter2_first":   # one pass that also writes the first iteration's V, then the batch on both
        assert niter == 2
        Vf = chmy.VectorField(arch, g, dt)
        for f, v in zip(Vf, sets[0].V):
            f.data.copy_(v.data)
        a, b = sets[0], sets[1]
        _lib.call("chmy_stokes_iter2_first", C.byref(g.desc()), _arr(list(a.V)), _arr(list(a.tau)),
                  C.byref(a.Pr.desc()), _arr(list(b.V)), _arr(list(b.tau)), C.byref(b.Pr.desc()), _arr(list(Vf)),
                  _arr(list(told)), C.byref(eta.desc()) if eta is not None else None,
                  C.byref(eml.desc()) if eml is not None else None, C.byref(rg.desc()) if rg is not None else None,
                  counts, comp, conds, flags, *scal, chmy.current_stream())
        for V in (b.V, Vf):
            _lib.call("chmy_bc_batchset", C.byref(g.desc()), counts, _arr([V[c] for c in comp]), conds,
                      chmy.current_stream())
        chmy.synchronize()
        return [f.data.cpu().numpy() for f in b.fields()], [f.data.cpu().numpy() for f in Vf]
    if mode == "iter2":   # one pass per two iterations, the batch after each pass
        assert niter % 2 == 0
        for _ in range(niter // 2):
            a, b = sets[par.value], sets[1 - par.value]
            a_ = args(a, b)[:13]   # the descriptors without divV, r_V
            _lib.call("chmy_stokes_iter2", *a_, counts, comp, conds, flags, *scal, chmy.current_stream())
            _lib.call("chmy_bc_batchset", C.byref(g.desc()), counts, _arr([b.V[c] for c in comp]), conds,
                      chmy.current_stream())
            par.value = 1 - par.value
    elif mode == "single-flag":   # chmy_stokes_iterate told to keep single iterations
        fs = flags | ITER_SINGLE
        _lib.call("chmy_stokes_iterate", *args(sets[0], sets[1]), counts, comp, conds, niter, fs, fs, *scal,
                  C.byref(par), chmy.current_stream())
    elif pairs:   # free-slip batches run in pairs by default (ITER_PAIRS is accepted and ignored)
        fp = flags | (ITER_PAIRS if niter % 2 else 0)
        _lib.call("chmy_stokes_iterate", *args(sets[0], sets[1]), counts, comp, conds, niter, fp, fp, *scal,
                  C.byref(par), chmy.current_stream())
    else:
        for _ in range(niter):
            a, b = sets[par.value], sets[1 - par.value]
            _lib.call("chmy_stokes_iter", *args(a, b), flags, *scal, chmy.current_stream())
            _lib.call("chmy_bc_batchset", C.byref(g.desc()), counts, _arr([b.V[c] for c in comp]), conds,
                      chmy.current_stream())
            par.value = 1 - par.value
    chmy.synchronize()
    return [f.data.cpu().numpy() for f in sets[par.value].fields()]


SHEAR = ((1.0, -1.0), (0.0, 0.0), (0.0, 0.0), (-1.0, 1.0))
GENERAL = ((0.7, -1.3), (0.25, -0.6), (-0.4, 0.9), (0.3, -0.2))


def _compare(a, b, what):
    names = ["Vx", "Vy", "txx", "tyy", "txy", "Pr"] if len(a) == 6 else \
        ["Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "Pr"]
    for n, x, y in zip(names, a, b):
        bits = np.int64 if x.dtype == np.float64 else np.int32
        if not np.array_equal(x.view(bits), y.view(bits)):
            d = np.nanmax(np.abs(x.astype(np.float64) - y.astype(np.float64)))
            bad = np.argwhere(x.view(bits) != y.view(bits))
            raise AssertionError(f"{what}: {n} differs (max |d| {d:.3e}, {len(bad)} points, first {bad[:4].tolist()})")


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("visc", ["field", "scalar"])
@pytest.mark.parametrize("dims", [(129, 23), (70, 45), (118, 30), (5, 3), (2, 2)])
def test_pair_pass_equals_single_iterations(arch, T, exact, visc, dims):
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=dims)
    for vals in (SHEAR, GENERAL):
        for niter in (2, 5):
            a = _run(arch, g, T, exact, visc, None, vals, niter, True, 11)
            b = _run(arch, g, T, exact, visc, None, vals, niter, False, 11)
            _compare(a, b, f"dims {dims} vals {vals} niter {niter}")


@pytest.mark.parametrize("T,forcing", [(np.float64, "same"), (np.float32, "same"), (np.float32, "bf16")])
@pytest.mark.parametrize("exact", [True, False])
def test_pair_pass_with_correction_forcing(arch, T, forcing, exact):
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=(70, 45))
    a = _run(arch, g, T, exact, "field", forcing, GENERAL, 4, True, 12)
    b = _run(arch, g, T, exact, "field", forcing, GENERAL, 4, False, 12)
    _compare(a, b, f"forcing {forcing}")


SHEAR3 = [[[1.0, -1.0], [0.0, 0.0], [0.0, 0.0]],
          [[0.0, 0.0], [0.0, 0.0], [0.0, 0.0]],
          [[0.0, 0.0], [0.0, 0.0], [-1.0, 1.0]]]
GENERAL3 = [[[0.7, -1.3], [0.25, -0.6], [0.1, 0.35]],
            [[-0.4, 0.9], [0.3, -0.2], [-0.15, 0.05]],
            [[0.2, 0.45], [-0.3, 0.6], [0.8, -0.5]]]


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("visc", ["field", "scalar"])
@pytest.mark.parametrize("dims", [(61, 13, 17), (20, 14, 35), (130, 30, 41), (5, 4, 3), (2, 2, 2), (59, 13, 11),
                                  (60, 10, 9), (118, 20, 9), (40, 40, 7), (15, 40, 9), (75, 13, 9), (76, 70, 12),
                                  (86, 20, 7)])
def test_pair_pass_3d_equals_single_iterations(arch, T, exact, visc, dims):
    """3D: x spans up to three 60-column tiles, y several 9-row tiles, z
    several chunks (the host picks chunks of 2 to 6 planes at these sizes,
    37 at 255^3), down to the smallest grid the pass takes; 59 and 60 cells
    put the right x-ghost column on its tile's first and second owned lane,
    118 fills two tiles exactly, and 15 to 86 leave 17 to 42 columns in the
    last tile."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=dims)
    for vals in (SHEAR3, GENERAL3):
        for niter in (2, 3):
            a = _run(arch, g, T, exact, visc, None, vals, niter, True, 21)
            b = _run(arch, g, T, exact, visc, None, vals, niter, False, 21)
            _compare(a, b, f"dims {dims} niter {niter}")


@pytest.mark.parametrize("T,forcing", [(np.float64, "same"), (np.float32, "same"), (np.float32, "bf16")])
@pytest.mark.parametrize("exact", [True, False])
def test_pair_pass_3d_with_correction_forcing(arch, T, forcing, exact):
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=(61, 13, 17))
    a = _run(arch, g, T, exact, "field", forcing, GENERAL3, 4, True, 22)
    b = _run(arch, g, T, exact, "field", forcing, GENERAL3, 4, False, 22)
    _compare(a, b, f"forcing {forcing}")


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("forcing", [None, "same"])
def test_pair_pass_2d_with_rho_g(arch, T, exact, forcing):
    """rho g (examples/stokes_2d_inc_ve_T.jl passes it to update_velocity!)
    with and without correction forcing."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=(70, 45))
    a = _run(arch, g, T, exact, "field", forcing, GENERAL, 5, True, 31, rhog=True)
    b = _run(arch, g, T, exact, "field", forcing, GENERAL, 5, False, 31, rhog=True)
    _compare(a, b, f"rho g, forcing {forcing}")


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("forcing", [None, "same"])
def test_pair_pass_3d_with_rho_g(arch, T, exact, forcing):
    """rho g at Vz (examples/stokes_3d_inc_ve_T.jl passes rho g_z to
    update_velocity!); with forcing as well the call keeps single iterations."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=(61, 13, 17))
    a = _run(arch, g, T, exact, "field", forcing, GENERAL3, 4, True, 32, rhog=True)
    b = _run(arch, g, T, exact, "field", forcing, GENERAL3, 4, False, 32, rhog=True)
    _compare(a, b, f"3D rho g, forcing {forcing}")


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("visc", ["field", "scalar"])
def test_iter2_equals_two_single_iterations(arch, nd, T, visc):
    """chmy_stokes_iter2 + the batch, twice, against four single iterations."""
    if nd == 2:
        g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=(70, 45))
        vals = GENERAL
    else:
        g = chmy.UniformGrid(arch, origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=(61, 13, 17))
        vals = GENERAL3
    for exact in (True, False):
        a = _run(arch, g, T, exact, visc, None, vals, 4, True, 41, rhog=True, mode="iter2")
        b = _run(arch, g, T, exact, visc, None, vals, 4, False, 41, rhog=True)
        _compare(a, b, f"iter2 {nd}D exact {exact}")


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("visc", ["field", "scalar"])
@pytest.mark.parametrize("rhog", [True, False])
@pytest.mark.parametrize("dims", [(61, 13, 17), (2, 2, 2), (59, 13, 11), (130, 30, 41)])
def test_iter2_first_keeps_the_first_iteration(arch, exact, visc, rhog, dims):
    """chmy_stokes_iter2_first + the batch on both outputs: the pass's state
    equals two single iterations and V_first equals V after the first."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=dims)
    for vals in (SHEAR3, GENERAL3):
        state, vf = _run(arch, g, np.float64, exact, visc, None, vals, 2, True, 51, rhog=rhog, mode="iter2_first")
        _compare(state, _run(arch, g, np.float64, exact, visc, None, vals, 2, False, 51, rhog=rhog),
                 f"iter2_first state, dims {dims}")
        one = _run(arch, g, np.float64, exact, visc, None, vals, 1, False, 51, rhog=rhog)
        for n, x, y in zip(("Vx", "Vy", "Vz"), vf, one[:3]):
            if not np.array_equal(x.view(np.int64), y.view(np.int64)):
                bad = np.argwhere(x.view(np.int64) != y.view(np.int64))
                raise AssertionError(f"V_first {n} differs at {len(bad)} points, first {bad[:4].tolist()}")


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("visc", ["field", "scalar"])
@pytest.mark.parametrize("rhog", [True, False])
@pytest.mark.parametrize("dims", [(70, 45), (2, 2), (129, 23)])
def test_iter2_first_2d_keeps_the_first_iteration(arch, exact, visc, rhog, dims):
    """2D: the same as test_iter2_first_keeps_the_first_iteration."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=dims)
    for vals in (SHEAR, GENERAL):
        state, vf = _run(arch, g, np.float64, exact, visc, None, vals, 2, True, 52, rhog=rhog, mode="iter2_first")
        _compare(state, _run(arch, g, np.float64, exact, visc, None, vals, 2, False, 52, rhog=rhog),
                 f"iter2_first state, dims {dims}")
        one = _run(arch, g, np.float64, exact, visc, None, vals, 1, False, 52, rhog=rhog)
        for n, x, y in zip(("Vx", "Vy"), vf, one[:2]):
            if not np.array_equal(x.view(np.int64), y.view(np.int64)):
                bad = np.argwhere(x.view(np.int64) != y.view(np.int64))
                raise AssertionError(f"V_first {n} differs at {len(bad)} points, first {bad[:4].tolist()}")


def test_iter2_first_rejects_what_it_does_not_run(arch):
    """An FP32 pass, an FP32 V_first beside FP64 fields, or V_first aliasing the
    input V, is refused before any launch."""
    g = chmy.UniformGrid(arch, origin=(-0.5,) * 3, extent=(1.0,) * 3, dims=(8, 8, 8))
    counts, comp, conds = _batch3(GENERAL3)
    p = pt_params(g)
    scal = (1.0, 0.5, 1.0, 0.3, p["dtau_Pr"], p["dtau_r"], p["nudtau"])

    def call(dt, Vf_dt=None, alias=False):
        a, b = _Set(arch, g, dt), _Set(arch, g, dt)
        Vf = a.V if alias else chmy.VectorField(arch, g, Vf_dt or dt)
        told = chmy.TensorField(arch, g, dt)
        return _lib.lib().chmy_stokes_iter2_first(
            C.byref(g.desc()), _arr(list(a.V)), _arr(list(a.tau)), C.byref(a.Pr.desc()), _arr(list(b.V)),
            _arr(list(b.tau)), C.byref(b.Pr.desc()), _arr(list(Vf)), _arr(list(told)), None, None, None,
            counts, comp, conds, 0, *scal, chmy.current_stream())

    assert call(torch.float32) != 0
    assert call(torch.float64, torch.float32) != 0
    assert call(torch.float64, alias=True) != 0   # V_first must not be the input V
    assert call(torch.float64) == 0
    chmy.synchronize()


def test_single_flag_keeps_results(arch):
    """CHMY_ITER_SINGLE: chmy_stokes_iterate one iteration per launch, same bits."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=(20, 14, 35))
    a = _run(arch, g, np.float64, False, "field", None, GENERAL3, 4, True, 43, mode="single-flag")
    b = _run(arch, g, np.float64, False, "field", None, GENERAL3, 4, False, 43)
    _compare(a, b, "single flag")


def test_iter2_rejects_other_batches(arch):
    """A V batch that is not the free-slip batch is refused on the host."""
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=(8, 6))
    sets = [_Set(arch, g, torch.float64), _Set(arch, g, torch.float64)]
    told = chmy.TensorField(arch, g, torch.float64)
    D, N, S = _lib.DIRICHLET, _lib.NEUMANN, _lib.VALUE_SCALAR
    rows = [(0, N, 0.0), (1, N, 0.0), (0, D, 0.0), (1, N, 0.0), (0, N, 0.0), (1, D, 0.0), (0, N, 0.0), (1, D, 0.0)]
    counts = (C.c_int32 * 6)(2, 2, 2, 2, 0, 0)
    comp = (C.c_int32 * 8)(*[c for c, _, _ in rows])
    conds = (_lib.ChmyBC * 8)(*[_lib.ChmyBC(k, S, v, None) for _, k, v in rows])
    p = pt_params(g)
    with pytest.raises(_lib.ChmyError, match="free-slip"):
        _lib.call("chmy_stokes_iter2", C.byref(g.desc()), _arr(list(sets[0].V)), _arr(list(sets[0].tau)),
                  C.byref(sets[0].Pr.desc()), _arr(list(sets[1].V)), _arr(list(sets[1].tau)),
                  C.byref(sets[1].Pr.desc()), _arr(list(told)), None, None, None, None, None, counts, comp, conds,
                  0, 0.8, 0.5, 1.0, 0.37, p["dtau_Pr"], p["dtau_r"], p["nudtau"], chmy.current_stream())
