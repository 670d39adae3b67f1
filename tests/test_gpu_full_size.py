"""Parity at the workloads' full sizes, through properties that do not depend
on size: two formulations of the same iteration give the same bits.

- 3D PT Stokes at 255³ (spherical inclusion, η = 1e-3 inside, times exp(-T)),
  FP64: the fused exact iteration (one launch + the V batch) against the
  reference's update_stress! + update_velocity! kernels and the V batch
  (stokes_3d_inc_ve_T.jl:164-165), which tests/test_gpu_kernels.py pins to the
  oracle point by point; and the two-iteration pass against single fused
  iterations, fast form.
- 2D PT Stokes at 2047² (32 inclusions, log10 contrast in [-3, 3]), FP64 and
  FP32: the two-iteration pass with the in-kernel free-slip batch against
  single fused iterations with the batch launched between them.

Each compares the whole padded storage bitwise on the device."""
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
