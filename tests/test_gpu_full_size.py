"""Parity at the workloads' full sizes: one exact fused iteration at 255³
against the oracle on seeded random storage, and properties that do not
depend on size (two formulations of the same iteration give the same bits):

- 3D PT Stokes at 255³ (spherical inclusion, η = 1e-3 inside, times exp(-T)),
  FP64: the fused exact iteration (one launch + the V batch) against the
  reference's update_stress! + update_velocity! kernels and the V batch
  (stokes_3d_inc_ve_T.jl:164-165), which tests/test_gpu_kernels.py pins to the
  oracle point by point; and the two-iteration pass against single fused
  iterations, fast form.
- 2D PT Stokes at 2047² (32 inclusions, log10 contrast in [-3, 3]), FP64 and
  FP32: the two-iteration pass with the in-kernel free-slip batch against
  single fused iterations with the batch launched between them.

Each compares the whole padded storage bitwise."""
import pytest
import torch

import chmy
from chmy import configs
from chmy.stokes import PTStokes, pure_shear_initial, pure_shear_values

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _state(S):
    return list(S.V) + list(S.tau) + [S.Pr]


def _bitwise(a, b, what):
    bits = torch.int64 if a.data.dtype == torch.float64 else torch.int32
    x, y = a.data.view(bits), b.data.view(bits)
    if not torch.equal(x, y):
        n = int((x != y).sum())
        d = float((a.data.double() - b.data.double()).abs().max())
        raise AssertionError(f"{what}: {n} points differ (max |d| {d:.3e})")


def _solver(arch, g, eta, dt, nd, dtype=torch.float64, exact=False):
    S = PTStokes(arch, g, dtype=dtype, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(nd),
                 exact=exact)
    pure_shear_initial(g, S.V)
    S.apply_bc()
    S.sync_buffers()
    return S


def test_3d_255_fused_exact_equals_reference_kernels(arch):
    g, eta, dt = configs.stokes3d_inclusion(arch, 255)
    A = _solver(arch, g, eta, dt, 3, exact=True)
    B = _solver(arch, g, eta, dt, 3, exact=True)
    for _ in range(3):
        A.iteration()
        B.iteration_unfused()
    chmy.synchronize()
    assert float(A.V[0].data.abs().max()) > 0 and float(A.tau[3].data.abs().max()) > 0
    for k, (a, b) in enumerate(zip(_state(A), _state(B))):
        _bitwise(a, b, f"255^3 fused vs reference kernels, field {k}")
    del A, B
    torch.cuda.empty_cache()


def test_3d_255_pairs_equal_single_iterations(arch):
    g, eta, dt = configs.stokes3d_inclusion(arch, 255)
    A = _solver(arch, g, eta, dt, 3)
    B = _solver(arch, g, eta, dt, 3)
    A.iterate(4, pairs=True)
    for _ in range(4):
        B.iteration()
    chmy.synchronize()
    for k, (a, b) in enumerate(zip(_state(A), _state(B))):
        _bitwise(a, b, f"255^3 pairs vs single, field {k}")
    del A, B
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_2d_2047_pairs_equal_single_iterations(arch, dtype):
    g, eta, dt = configs.stokes2d_inclusions(arch, 2047)
    A = _solver(arch, g, eta, dt, 2, dtype=dtype)
    B = _solver(arch, g, eta, dt, 2, dtype=dtype)
    A.iterate(6)   # 2D: three two-iteration passes
    for _ in range(6):
        B.iteration()
    chmy.synchronize()
    assert float(A.V[1].data.abs().max()) > 0
    for k, (a, b) in enumerate(zip(_state(A), _state(B))):
        _bitwise(a, b, f"2047^2 {dtype} pairs vs single, field {k}")


def test_3d_255_correction_sweep_pairs_equal_single(arch):
    """The FP32 defect-correction sweep at 255³ (bf16 forcing from the f64
    residual of the initial state): two-iteration passes against single fused
    iterations."""
    from chmy.stokes import MixedPTStokes
    g, eta, dt = configs.stokes3d_inclusion(arch, 255)
    Ms = []
    for _ in range(2):
        M = MixedPTStokes(arch, g, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(3), inner=6)
        pure_shear_initial(g, M.x.V)
        M.x.apply_bc()
        M.refresh()
        M.reset_correction()
        Ms.append(M)
    Ms[0].d.iterate(6, pairs=True)
    for _ in range(6):
        Ms[1].d.iteration()
    chmy.synchronize()
    assert float(Ms[0].d.V[0].data.abs().max()) > 0
    for k, (a, b) in enumerate(zip(_state(Ms[0].d), _state(Ms[1].d))):
        _bitwise(a, b, f"255^3 FP32 correction sweep pairs vs single, field {k}")


def test_3d_255_fused_exact_equals_oracle(arch):
    """One fused exact iteration at 255³ FP64 on seeded random storage (every
    field, η and η_ml in [1e-3, 3]) against the oracle's update_stress! and
    update_velocity! on the same storage: V, τ, Pr, ∇V and r_V bitwise."""
    import ctypes as C

    import numpy as np

    import oracle as O
    from _helpers import assert_same, ogrid, pair, vloc
    from chmy import Center as Cc
    from chmy import _lib
    from chmy.kernels import _arr
    g = chmy.UniformGrid(arch, origin=(-0.5,) * 3, extent=(1.0,) * 3, dims=(255,) * 3)
    og = ogrid(g)
    rng = np.random.default_rng(255)
    T = np.float64
    tl = list(chmy.tensor_locations(3).values())
    V = [pair(arch, g, og, vloc(D, 3), T, rng=rng) for D in range(3)]
    tau = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    Pr = pair(arch, g, og, Cc(), T, rng=rng)
    told = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    eta = pair(arch, g, og, Cc(), T, rng=rng, lo=1e-3, hi=3.0)
    eml = pair(arch, g, og, Cc(), T, rng=rng, lo=1e-3, hi=3.0)
    rg = pair(arch, g, og, vloc(2, 3), T, rng=rng)
    dV = pair(arch, g, og, Cc(), T, fill=0.0)
    rV = [pair(arch, g, og, vloc(D, 3), T, fill=0.0) for D in range(3)]
    Vo = [chmy.Field(arch, g, f.loc, f.dtype) for f, _ in V]
    tauo = [chmy.Field(arch, g, f.loc, f.dtype) for f, _ in tau]
    Pro = chmy.Field(arch, g, Cc(), Pr[0].dtype)
    for a, (b, _) in zip(Vo + tauo + [Pro], V + tau + [Pr]):
        a.data.copy_(b.data)
    G, dt, eta_s = 1.0, 0.37, 0.8
    eve = 1.0 / (1.0 / eta_s + 1.0 / (G * dt))
    p = O.pt_params(og)
    _lib.call("chmy_stokes_iter", C.byref(g.desc()), _arr([a for a, _ in V]), _arr([a for a, _ in tau]),
              C.byref(Pr[0].desc()), _arr(Vo), _arr(tauo), C.byref(Pro.desc()), _arr([a for a, _ in told]),
              C.byref(eta[0].desc()), C.byref(eml[0].desc()), C.byref(rg[0].desc()), None, None,
              C.byref(dV[0].desc()), _arr([a for a, _ in rV]), 1 | 2, eta_s, eve, G, dt,
              p["dtau_Pr"], p["dtau_r"], p["nudtau"], chmy.current_stream())
    chmy.synchronize()
    H = lambda L: [b for _, b in L]
    O.update_stress(og, H(tau), Pr[1], dV[1], H(V), H(told), eta[1], eve, G, dt, p["dtau_Pr"], p["dtau_r"])
    O.update_velocity(og, H(V), H(rV), Pr[1], H(tau), rg[1], eve, p["nudtau"], eta_ml=eml[1], G=G, dt=dt)
    for D, (f, (_, h)) in enumerate(zip(Vo, V)):
        assert_same(f, h, f"V[{D}]")
    for c, (f, (_, h)) in enumerate(zip(tauo, tau)):
        assert_same(f, h, f"tau[{c}]")
    assert_same(Pro, Pr[1], "Pr")
    assert_same(dV[0], dV[1], "divV")
    for D in range(3):
        assert_same(rV[D][0], rV[D][1], f"r_V[{D}]")


def test_3d_255_fast_form_tracks_exact_form(arch):
    """200 PT iterations at 255³ FP64 (inclusion workload, V batch between
    iterations): the fast form (7 divisions per point) stays within 1e-11 of
    the exact form relative to each field's max|value|."""
    g, eta, dt = configs.stokes3d_inclusion(arch, 255)
    A = _solver(arch, g, eta, dt, 3, exact=True)
    B = _solver(arch, g, eta, dt, 3, exact=False)
    A.iterate(200)
    B.iterate(200)
    chmy.synchronize()
    for k, (a, b) in enumerate(zip(_state(A), _state(B))):
        scale = float(a.data.abs().max())
        assert scale > 0, k
        d = float((a.data - b.data).abs().max())
        assert d <= 1e-11 * scale, (k, d, scale)
