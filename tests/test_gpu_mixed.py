"""FP32 PT sweeps with FP64 defect correction (build extension).

- chmy_stokes_residual equals the oracle's residual restatement bitwise (f64
  norms fields) and its f64/f32 forcing is G·dt·R_tau, R_P, R_V;
- the defect-correction solve reaches the same f64 fixed point as the plain
  f64 PT solve: both converged to 1e-10, fields agree to ~1e-8 relative, and
  the final f64 residual of the mixed solve is below the tolerance.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy import _lib
from chmy.kernels import _arr
from chmy.stokes import MixedPTStokes, PTStokes, inclusion_setup_3d, pure_shear_initial, pure_shear_values
import oracle as O
from _helpers import assert_same, dev, ogrid, pair, vloc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _bf16_bits(a):
    """Round-to-nearest-even bf16 bits of float32(a) (numpy), the device rule."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _devbits(f):
    return f.data.detach().view(torch.int16).cpu().numpy().view(np.uint16)


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("visc", ["scalar", "field"])
@pytest.mark.parametrize("out", [np.float64, np.float32, "bf16"])
def test_residual_kernel(arch, nd, visc, out):
    g = chmy.UniformGrid(arch, origin=(-0.5,) * nd, extent=(1.0,) * nd, dims=(37, 12, 9)[:nd] if nd == 3 else (70, 33))
    og = ogrid(g)
    rng = np.random.default_rng(70 + nd)
    T = np.float64
    tl = list(chmy.tensor_locations(nd).values())
    V = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    tau = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    told = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    Pr = pair(arch, g, og, Cc(), T, rng=rng)
    eta = pair(arch, g, og, Cc(), T, rng=rng, lo=1e-2, hi=2.0)
    rg = pair(arch, g, og, vloc(nd - 1, nd), T, rng=rng)
    dV = pair(arch, g, og, Cc(), T, fill=0)
    rV = [pair(arch, g, og, vloc(D, nd), T, fill=0) for D in range(nd)]
    to_t = {np.float64: torch.float64, np.float32: torch.float32, "bf16": torch.bfloat16}[out]
    toc = [chmy.Field(arch, g, l, to_t) for l in tl]
    fV = [chmy.Field(arch, g, vloc(D, nd), to_t) for D in range(nd)]
    fP = chmy.Field(arch, g, Cc(), to_t)
    fv = visc == "field"
    G, dt, eta_s = 1.0, 0.3, 0.7
    _lib.call("chmy_stokes_residual", C.byref(g.desc()), _arr([a for a, _ in V]), _arr([a for a, _ in tau]),
              C.byref(Pr[0].desc()), _arr([a for a, _ in told]), C.byref(eta[0].desc()) if fv else None,
              C.byref(rg[0].desc()), None, eta_s, G, dt, _arr(toc), _arr(fV), C.byref(fP.desc()), C.byref(dV[0].desc()),
              _arr([a for a, _ in rV]), chmy.current_stream())
    chmy.synchronize()
    H = lambda L: [b for _, b in L]
    Rt = [O.HField(og, tuple(l.code for l in lc)) for lc in tl]
    RV = [O.HField(og, tuple(l.code for l in vloc(D, nd))) for D in range(nd)]
    RP = O.HField(og, 0)
    O.stokes_residual(og, H(V), H(tau), Pr[1], H(told), eta[1] if fv else eta_s, rg[1], G, dt, Rt, RV, RP)
    sl = tuple(slice(1, -1) for _ in range(nd))
    if out == "bf16":   # f64 value -> f32 -> bf16, both rounded to nearest even
        same = lambda f, ref: np.array_equal(_devbits(f)[sl], _bf16_bits(ref[sl]))
    else:
        same = lambda f, ref: np.array_equal(dev(f)[sl], ref[sl].astype(out))
    for D in range(nd):
        assert np.array_equal(dev(rV[D][0])[sl], RV[D].arr[sl])
        assert same(fV[D], RV[D].arr)
    assert np.array_equal(dev(dV[0])[sl], -RP.arr[sl])
    assert same(fP, RP.arr)
    for c in range(len(tl)):
        assert same(toc[c], (G * dt) * Rt[c].arr)


def _setup_2d(arch, n):
    g = chmy.UniformGrid(arch, origin=(-0.5, -0.5), extent=(1.0, 1.0), dims=(n, n))
    xs = g.coords(Cc(), 0, first=-1, count=n + 4)
    X, Y = xs[None, :], xs[:, None]
    eta = np.where(X ** 2 + Y ** 2 < 0.1 ** 2, 1e-2, 1.0)
    return g, eta


@pytest.mark.parametrize("nd", [2, 3])
def test_defect_correction_reaches_f64_fixed_point(arch, nd):
    n = 48 if nd == 2 else 24
    if nd == 2:
        g, eta = _setup_2d(arch, n)
    else:
        g = chmy.UniformGrid(arch, origin=(-0.5,) * 3, extent=(1.0,) * 3, dims=(n,) * 3)
        eta = inclusion_setup_3d(g, eta_incl=1e-2, radius=0.2)
    G, dt, eps = 1.0, 0.25, 1e-10
    ref = PTStokes(arch, g, eta_field=eta, G=G, dt=dt, normal_values=pure_shear_values(nd))
    pure_shear_initial(g, ref.V)
    ref.apply_bc()
    ref.sync_buffers()
    it64, err64 = ref.solve(eps=eps, max_iter=200000, ncheck=n)
    assert max(err64) < eps, err64
    # the f64 PT fixed point also satisfies the stress equation
    assert max(ref.true_residual()) < 1e-8
    mx = MixedPTStokes(arch, g, eta_field=eta, G=G, dt=dt, normal_values=pure_shear_values(nd), inner=2 * n)
    pure_shear_initial(g, mx.x.V)
    mx.x.apply_bc()
    sweeps, err = mx.solve(eps=eps, max_cycles=2000)
    assert max(err) < eps, err
    # both sit within the tolerance of the same discrete solution
    for a, b in zip(list(ref.V) + [ref.Pr], list(mx.x.V) + [mx.x.Pr]):
        ai, bi = a.interior().cpu().numpy(), b.interior().cpu().numpy()
        scale = max(np.max(np.abs(ai)), 1.0)
        assert np.max(np.abs(ai - bi)) / scale < 1e-6
    # the f32 sweeps did not need vastly more iterations than the f64 solve
    assert sweeps < 3 * it64 + 10 * mx.inner


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("exact", [False, True])
def test_bf16_forcing_sweep_equals_f32_forcing_of_same_values(arch, nd, exact):
    """An FP32 sweep reading bf16 tau_old/f_V/f_P gives the same bits as one
    reading FP32 fields holding the same (bf16-representable) values."""
    dims = (40, 19, 11)[:nd] if nd == 3 else (70, 45)
    g = chmy.UniformGrid(arch, origin=(-0.5,) * nd, extent=(1.0,) * nd, dims=dims)
    rng = np.random.default_rng(9)
    eta = np.exp(rng.uniform(-1, 1, tuple(n + 4 for n in reversed(dims))))
    runs = []
    for fdt in (torch.bfloat16, torch.float32):
        S = PTStokes(arch, g, dtype=torch.float32, eta_field=eta, G=1.0, dt=0.3, exact=exact)
        S.tau_old = chmy.TensorField(arch, g, fdt)
        S.f_V = chmy.VectorField(arch, g, fdt)
        S.f_P = chmy.Field(arch, g, Cc(), fdt)
        r2 = np.random.default_rng(3)
        for f in list(S.V) + list(S.tau) + [S.Pr]:
            f.data.copy_(torch.from_numpy(r2.uniform(-1, 1, tuple(f.data.shape)).astype(np.float32)))
        for f in list(S.tau_old) + list(S.f_V) + [S.f_P]:
            v = torch.from_numpy(r2.uniform(-1, 1, tuple(f.data.shape)).astype(np.float32))
            f.data.copy_(v.to(torch.bfloat16).to(fdt))
        S.sync_buffers()
        for _ in range(3):
            S.iteration()
        runs.append([f.data.cpu().numpy() for f in list(S.V) + list(S.tau) + [S.Pr]])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
