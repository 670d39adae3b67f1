"""SIA forward kernels against the oracle (bitwise) and the hand-written
adjoint against finite differences of the GPU forward map."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy import _lib
from chmy.kernels import _arr
from chmy.sia import SIA, bed_bumps, smooth_field
import oracle as O
from _helpers import assert_same, dev, ogrid, pair, vloc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _setup(arch, n=(40, 33), seed=0, with_A=True):
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=n)
    og = ogrid(g)
    rng = np.random.default_rng(seed)
    H, hH = pair(arch, g, og, Cc(), fill=0)
    xi, yi = g.coords(Cc(), 0), g.coords(Cc(), 1)
    hH.set_interior(0.3 + 0.2 * np.exp(-(xi[:, None] ** 2 + yi[None, :] ** 2) / 0.3) + 0.05 * rng.random(n))
    O.bc_apply(og, [(hH, O.Neumann())])
    chmy.set_storage(H, hH.arr)
    B, hB = pair(arch, g, og, Cc(), fill=0)
    xs = g.coords(Cc(), 0, first=-1, count=n[0] + 4)
    ys = g.coords(Cc(), 1, first=-1, count=n[1] + 4)
    hB.xyz()[...] = np.exp(-(xs[:, None] / 0.5) ** 2 - (ys[None, :] / 0.5) ** 2)
    chmy.set_storage(B, hB.arr)
    A = hA = None
    if with_A:
        A, hA = pair(arch, g, og, Cc(), rng=rng, lo=0.5, hi=1.5)
    return g, og, H, hH, B, hB, A, hA


@pytest.mark.parametrize("with_A", [False, True])
def test_sia_flux_update_bitwise(arch, with_A):
    g, og, H, hH, B, hB, A, hA = _setup(arch, with_A=with_A)
    q = [chmy.Field(arch, g, vloc(D, 2)) for D in range(2)]
    hq = [O.HField(og, tuple(l.code for l in vloc(D, 2))) for D in range(2)]
    As, n, mb, dt = 0.5, 3, (0.5, 0.3, 0.25), 2e-3
    _lib.call("chmy_sia_flux", C.byref(g.desc()), _arr(q), C.byref(H.desc()), C.byref(B.desc()),
              C.byref(A.desc()) if with_A else None, As, n, chmy.current_stream())
    O.sia_flux(og, hq, hH, hB, A=hA, As=As, npow=n, mb=mb)
    for f, h in zip(q, hq):
        assert_same(f, h, "q")
    _lib.call("chmy_sia_update", C.byref(g.desc()), C.byref(H.desc()), C.byref(H.desc()), _arr(q), C.byref(B.desc()),
              *mb, dt, chmy.current_stream())
    Hn = hH.copy()
    O.sia_update(og, hH, Hn, hq, hB, dt, As=As, npow=n, mb=mb)
    assert_same(H, Hn, "H after update_H!")


def test_sia_fused_step_equals_reference_sequence(arch):
    """chmy_sia_step + Neumann == update_q!, q Dirichlet (all sides, sia.jl:72-73),
    update_H!, H Neumann (sia.jl:80-81), for 20 steps, bitwise."""
    g, og, H, hH, B, hB, A, hA = _setup(arch, seed=3)
    S = SIA(arch, g, hB.arr, logA=np.log(hA.arr), As=0.5, npow=3, mb=(0.5, 0.3, 0.25))
    S.reset(hH.arr)
    hq = [O.HField(og, tuple(l.code for l in vloc(D, 2))) for D in range(2)]
    dt = 2e-3
    for _ in range(20):
        S.step(dt)
        S.H.swap(S.H2)
        O.sia_flux(og, hq, hH, hB, A=hA, As=0.5, npow=3)
        O.bc_apply(og, [(hq[0], O.Dirichlet()), (hq[1], O.Dirichlet())])
        Hn = hH.copy()
        O.sia_update(og, hH, Hn, hq, hB, dt, As=0.5, npow=3)
        O.bc_apply(og, [(Hn, O.Neumann())])
        hH = Hn
    assert_same(S.H, hH, "H after 20 fused steps", interior_only=True)
    # D and the time-step rule (sia.jl:76)
    dmax = S.diffusivity_max()
    assert dmax == O.sia_Dmax(og, hH, hB, A=hA, As=0.5, npow=3)


@pytest.mark.parametrize("npow", [3, 2, 4])
def test_sia_run_equals_step_and_neumann_sequence(arch, npow):
    """chmy_sia_run (step + Neumann batch in one tiled launch, Δt per step)
    gives the same bits as chmy_sia_step followed by the Neumann batch, on a
    grid that is not a multiple of the tile (45 x 21), for 15 steps; n = 3
    takes the tile kernel compiled for n = 3, other exponents the generic one."""
    n = (45, 21)
    g, og, H, hH, B, hB, A, hA = _setup(arch, n=n, seed=7)
    S = SIA(arch, g, hB.arr, logA=np.log(hA.arr), As=0.5, npow=npow, mb=(0.5, 0.3, 0.25))
    T = SIA(arch, g, hB.arr, logA=np.log(hA.arr), As=0.5, npow=npow, mb=(0.5, 0.3, 0.25))
    S.reset(hH.arr)
    T.reset(hH.arr)
    dts = [2e-3 * (1 + 0.1 * k) for k in range(15)]
    for dt in dts:
        S.step(dt)
        S.H.swap(S.H2)
    T.forward(15, dts)
    assert np.array_equal(dev(T.H), dev(S.H))
    # checkpointed run: same final state, checkpoints equal the stepped states
    T.reset(hH.arr)
    T.forward(15, dts, ncheck=4)
    assert np.array_equal(dev(T.H), dev(S.H))
    assert T.stride == 4 and len(T.ckpt) == 4
    S.reset(hH.arr)
    for k, dt in enumerate(dts):
        if k % 4 == 0:
            assert np.array_equal(dev(T.ckpt[k // 4]), dev(S.H)), k
        S.step(dt)
        S.H.swap(S.H2)


@pytest.mark.parametrize("npow", [3, 2, 4])
@pytest.mark.parametrize("n", [(45, 21), (64, 32)])
def test_sia_adjoint_tiled_equals_five_pass(arch, n, npow):
    """The tiled one-launch adjoint step and the five gather passes give the
    same bits (λ and the log A gradient), interior and tile edges included,
    for the n = 3 instantiation and the generic exponent."""
    g, og, H, hH, B, hB, A, hA = _setup(arch, n=n, seed=11)
    S = SIA(arch, g, hB.arr, logA=np.log(hA.arr), As=0.5, npow=npow, mb=(0.5, 0.3, 0.25))
    rng = np.random.default_rng(2)
    lam0 = torch.from_numpy(rng.standard_normal((n[1], n[0])).T.copy()).to(S.lam.data.device)
    out = []
    for passes in (1, 5):
        S.reset(hH.arr)
        S.grad.data.zero_()
        S.lam.data.zero_()
        S.lam2.data.zero_()
        S.lam.interior().copy_(lam0)
        S._adjoint_step(S.H, 2e-3, passes=passes)
        out.append((S.lam.interior().cpu().numpy().copy(), S.grad.interior().cpu().numpy().copy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_sia_reverse_sweep_equals_stepwise_adjoint(arch):
    """chmy_sia_reverse (checkpoints, segment recompute riding along the
    adjoint launches, a short last segment) equals storing every state and
    applying the adjoint step by step, bitwise."""
    n = (40, 33)
    g, og, H, hH, B, hB, A, hA = _setup(arch, n=n, seed=3)
    S = SIA(arch, g, hB.arr, logA=np.log(hA.arr), As=0.5, npow=3, mb=(0.5, 0.3, 0.25))
    nsteps = 23
    S.reset(hH.arr)
    dts = S.forward(nsteps)
    H_obs = S.H.interior().clone() * 0.97
    J, grad = S.gradient(hH.arr, H_obs, nsteps, dts, ncheck=5)   # stride 5, last segment 3 steps
    lam_rev = S.lam.interior().cpu().numpy().copy()
    # stepwise: every state stored
    S.reset(hH.arr)
    states = [S._field() for _ in range(nsteps + 1)]
    states[0].data.copy_(S.H.data)
    for k in range(nsteps):
        S.run(states[k:k + 2], dts[k:k + 1])
    S.grad.data.zero_()
    S.lam.data.zero_()
    S.lam2.data.zero_()
    S.lam.interior().copy_(states[-1].interior() - H_obs)
    for k in reversed(range(nsteps)):
        S._adjoint_step(states[k], dts[k])
    assert np.array_equal(S.lam.interior().cpu().numpy(), lam_rev)
    assert np.array_equal(S.grad.interior().cpu().numpy(), grad.cpu().numpy())


def test_sia_adjoint_step_equals_fd_jacobian_transpose(arch):
    """One adjoint step equals J^T λ with J the Jacobian of the GPU forward
    step + Neumann batch, by central differences on a 9x7 grid (state chosen
    away from the max/min kinks)."""
    n = (9, 7)
    g = chmy.UniformGrid(arch, origin=(0.0, 0.0), extent=(1.0, 0.8), dims=n)
    rng = np.random.default_rng(5)
    full = (n[1] + 4, n[0] + 4)
    B = 0.2 * rng.random(full)
    logA = np.log(rng.uniform(0.5, 1.5, full))
    S = SIA(arch, g, B, logA=logA, As=0.3, npow=3, mb=(0.05, -10.0, 100.0))
    H0 = np.zeros(full)
    H0[2:-2, 2:-2] = 1.0 + 0.5 * rng.random((n[1], n[0]))
    dt = 1e-3

    def F(h_int, la=None):
        S.set_logA(logA if la is None else la)
        H = H0.copy()
        H[2:-2, 2:-2] = h_int.reshape(n[1], n[0])
        S.reset(H)
        S.step(dt)
        return S.H2.interior().cpu().numpy().T.reshape(-1).copy()  # C-order over (y, x)

    h0 = H0[2:-2, 2:-2].reshape(-1).copy()
    m = h0.size
    eps = 1e-6
    Jh = np.zeros((m, m))
    for c in range(m):
        e = np.zeros(m)
        e[c] = eps
        Jh[:, c] = (F(h0 + e) - F(h0 - e)) / (2 * eps)
    Ja = np.zeros((m, m))
    for c in range(m):
        la = logA.copy()
        la[2 + c // n[0], 2 + c % n[0]] += eps
        lb = logA.copy()
        lb[2 + c // n[0], 2 + c % n[0]] -= eps
        Ja[:, c] = (F(h0, la) - F(h0, lb)) / (2 * eps)
    lam = rng.standard_normal(m)
    # adjoint
    S.set_logA(logA)
    H = H0.copy()
    S.reset(H)
    for passes in (1, 5):
        S.grad.data.zero_()
        S.lam.data.zero_()
        S.lam2.data.zero_()
        S.lam.interior().copy_(torch.from_numpy(lam.reshape(n[1], n[0]).T.copy()))
        S._adjoint_step(S.H, dt, passes=passes)
        lam_out = S.lam.interior().cpu().numpy().T.reshape(-1)
        gA = S.grad.interior().cpu().numpy().T.reshape(-1)
        assert np.allclose(lam_out, Jh.T @ lam, rtol=1e-6, atol=1e-7 * np.abs(Jh.T @ lam).max()), passes
        assert np.allclose(gA, Ja.T @ lam, rtol=1e-6, atol=1e-7 * np.abs(Ja.T @ lam).max()), passes


def test_sia_gradient_matches_central_differences(arch):
    """dJ/dlogA from the reverse sweep (uniform checkpoints, recomputed
    segments) matches central differences of J along random directions."""
    n = 64
    g = chmy.UniformGrid(arch, origin=(0.0, 0.0), extent=(250e3, 250e3), dims=(n, n))
    B = bed_bumps(g, seed=4, hmin=1500.0, hmax=2500.0)
    logA0 = np.full((n + 4, n + 4), math.log(2.8e-6))
    mb = (0.01, 1500.0, 2.0)
    S = SIA(arch, g, B, logA=logA0 + smooth_field(g, 0.3, seed=9), npow=3, mb=mb, dt_max=1.0)
    S.reset()
    nsteps = 60
    dts = S.forward(nsteps)
    H_obs = S.H.interior().clone()
    S.set_logA(logA0)
    J, grad = S.gradient(None, H_obs, nsteps, dts, ncheck=6)
    assert J > 0
    rng = np.random.default_rng(1)
    gn = grad.cpu().numpy()
    for _ in range(4):
        v = _interior_direction(rng, n)
        h = 1e-5
        S.set_logA(logA0 + h * v)
        S.reset()
        S.forward(nsteps, dts)
        Jp = S.misfit(H_obs)
        S.set_logA(logA0 - h * v)
        S.reset()
        S.forward(nsteps, dts)
        Jm = S.misfit(H_obs)
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.sum(gn * v[2:-2, 2:-2].T))
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def _interior_direction(rng, n):
    """A seeded Gaussian direction of log A on the cells only (zero on the
    halo and padding of the (n+4)^2 storage, which the gradient does not cover)."""
    v = np.zeros((n + 4, n + 4))
    v[2:-2, 2:-2] = rng.standard_normal((n, n))
    return v


def test_sia_gradient_at_the_configured_size(arch):
    """The gradient workload as configured: 512^2 cells, 300 fixed steps,
    30 checkpoints (stride 10, segments recomputed in the reverse sweep).
    dJ/dlogA along 8 seeded interior directions against central differences
    of the GPU forward J, h = 1e-5, to 1e-6 relative."""
    from chmy import configs
    g, B, la0, la1, mb, dt = configs.sia_ice_cap(arch, 512)
    nsteps, ncheck = 300, 30
    dts = [dt] * nsteps
    S = SIA(arch, g, B, logA=la1, As=0.0, npow=3, mb=mb)
    S.reset()
    S.forward(nsteps, dts)
    H_obs = S.H.interior().clone()
    S.set_logA(la0)
    J, grad = S.gradient(None, H_obs, nsteps, dts, ncheck)
    assert S.stride == 10 and len(S.ckpt) == 30
    assert J > 0
    gn = grad.cpu().numpy()

    def Jf(logA):
        S.set_logA(logA)
        S.reset()
        S.forward(nsteps, dts)
        return S.misfit(H_obs)

    rng = np.random.default_rng(16)
    h = 1e-5
    worst = 0.0
    for _ in range(8):
        v = _interior_direction(rng, 512)
        fd = (Jf(la0 + h * v) - Jf(la0 - h * v)) / (2 * h)
        ad = float(np.sum(gn * v[2:-2, 2:-2].T))
        worst = max(worst, abs(fd - ad) / abs(fd))
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    print(f"512^2, 300 steps, 30 checkpoints: worst relative FD error {worst:.2e}")


def test_sia_gradient_matches_cpu_reference(arch):
    """The GPU gradient (forward with checkpoints, reverse sweep) against the
    CPU reference: J and central differences of J from the oracle's forward
    (update_q!, q Dirichlet, update_H!, H Neumann; sia.jl:71-81) for the same
    Δt sequence, along random directions of log A."""
    n = 48
    g = chmy.UniformGrid(arch, origin=(0.0, 0.0), extent=(250e3, 250e3), dims=(n, n))
    og = ogrid(g)
    B = bed_bumps(g, seed=4, hmin=1500.0, hmax=2500.0)
    logA0 = np.full((n + 4, n + 4), math.log(2.8e-6))
    mb = (0.01, 1500.0, 2.0)
    S = SIA(arch, g, B, logA=logA0 + smooth_field(g, 0.3, seed=9), npow=3, mb=mb, dt_max=1.0)
    S.reset()
    nsteps = 40
    dts = S.forward(nsteps)
    H_obs = S.H.interior().cpu().numpy().copy()        # (x, y)
    S.set_logA(logA0)
    J, grad = S.gradient(None, S.H.interior().new_tensor(H_obs), nsteps, dts, ncheck=5)
    gn = grad.cpu().numpy()
    hB = O.HField(og, O.CENTER)
    hB.arr[...] = B

    def J_cpu(logA):
        hA = O.HField(og, O.CENTER)
        hA.arr[...] = np.exp(logA)
        hH = O.HField(og, O.CENTER)
        hq = [O.HField(og, tuple(l.code for l in vloc(D, 2))) for D in range(2)]
        for dt in dts:
            O.sia_flux(og, hq, hH, hB, A=hA, As=0.0, npow=3)
            O.bc_apply(og, [(hq[0], O.Dirichlet()), (hq[1], O.Dirichlet())])
            Hn = hH.copy()
            O.sia_update(og, hH, Hn, hq, hB, dt, As=0.0, npow=3, mb=mb)
            O.bc_apply(og, [(Hn, O.Neumann())])
            hH = Hn
        d = hH.arr[2:-2, 2:-2].T - H_obs
        return 0.5 * float(np.sum(d * d))

    Jc = J_cpu(logA0)
    assert J > 0 and abs(J - Jc) <= 1e-12 * Jc, (J, Jc)
    rng = np.random.default_rng(2)
    h = 1e-5
    for _ in range(4):
        v = _interior_direction(rng, n)
        fd = (J_cpu(logA0 + h * v) - J_cpu(logA0 - h * v)) / (2 * h)
        ad = float(np.sum(gn * v[2:-2, 2:-2].T))
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
