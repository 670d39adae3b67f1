"""The fused, double-buffered PT iteration (chmy_stokes_iter) equals the
reference's update_stress! + update_velocity! sequence bitwise, for one
iteration on random storage and for multi-iteration solves with the V batch."""
import math

import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy.stokes import PTStokes, inclusion_setup_3d, pure_shear_initial, pure_shear_values
import oracle as O
from _helpers import assert_same, dev, ogrid, pair, vloc

pytestmark = pytest.mark.gpu

TYPES = [np.float64, np.float32]


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _grid(nd, big_z=False):
    if nd == 2:
        return chmy.UniformGrid(origin=(-0.5, -0.5), extent=(1.0, 1.3), dims=(129, 23))
    if big_z:
        return chmy.UniformGrid(origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.0, 1.0), dims=(40, 14, 70))
    return chmy.UniformGrid(origin=(-0.5, -0.5, -0.5), extent=(1.0, 1.1, 0.9), dims=(67, 13, 9))


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("nd,big_z", [(2, False), (3, False), (3, True)])
@pytest.mark.parametrize("visc", ["scalar", "field"])
@pytest.mark.parametrize("wres", [0, 1])
@pytest.mark.parametrize("exact", [True, False])
def test_fused_iteration_equals_two_kernels(arch, T, nd, big_z, visc, wres, exact):
    """exact: bitwise; fast form: within a few ulp (relative 1e-13 f64 / 2e-5 f32)."""
    g = _grid(nd, big_z)
    og = ogrid(g)
    rng = np.random.default_rng(hash((nd, big_z, visc, wres)) % 2 ** 32)
    tl = list(chmy.tensor_locations(nd).values())
    V = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    tau = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    Pr = pair(arch, g, og, Cc(), T, rng=rng)
    told = [pair(arch, g, og, l, T, rng=rng) for l in tl]
    eta = pair(arch, g, og, Cc(), T, rng=rng, lo=1e-3, hi=3.0)
    eml = pair(arch, g, og, Cc(), T, rng=rng, lo=1e-3, hi=3.0)
    rg = pair(arch, g, og, vloc(nd - 1, nd), T, rng=rng)
    dV = pair(arch, g, og, Cc(), T, rng=rng)
    rV = [pair(arch, g, og, vloc(D, nd), T, rng=rng) for D in range(nd)]
    # output buffers start as copies of the inputs
    Vo = [chmy.Field(arch, g, f.loc, f.dtype) for f, _ in V]
    tauo = [chmy.Field(arch, g, f.loc, f.dtype) for f, _ in tau]
    Pro = chmy.Field(arch, g, Cc(), Pr[0].dtype)
    for a, (b, _) in zip(Vo + tauo + [Pro], V + tau + [Pr]):
        a.data.copy_(b.data)
    G, dt, eta_s = 1.0, 0.37, 0.8
    eve = 1.0 / (1.0 / eta_s + 1.0 / (G * dt))
    p = O.pt_params(og)
    import ctypes as C
    from chmy import _lib
    from chmy.kernels import _arr
    fv = visc == "field"
    # the kernel reads (V, tau, Pr) and writes only the output set
    inputs = [a for a, _ in V + tau] + [Pr[0]]
    before = [a.data.clone() for a in inputs]
    _lib.call("chmy_stokes_iter", C.byref(g.desc()), _arr([a for a, _ in V]), _arr([a for a, _ in tau]),
              C.byref(Pr[0].desc()), _arr(Vo), _arr(tauo), C.byref(Pro.desc()), _arr([a for a, _ in told]),
              C.byref(eta[0].desc()) if fv else None, C.byref(eml[0].desc()) if fv else None,
              C.byref(rg[0].desc()), None, None, C.byref(dV[0].desc()), _arr([a for a, _ in rV]),
              wres | (2 if exact else 0),
              eta_s, eve, G, dt,
              p["dtau_Pr"], p["dtau_r"], p["nudtau"], chmy.current_stream())
    chmy.synchronize()
    H = lambda L: [b for _, b in L]
    O.update_stress(og, H(tau), Pr[1], dV[1], H(V), H(told), eta[1] if fv else eta_s, eve, G, dt,
                    p["dtau_Pr"], p["dtau_r"])
    O.update_velocity(og, H(V), H(rV), Pr[1], H(tau), rg[1], eve, p["nudtau"], eta_ml=eml[1] if fv else None,
                      G=G, dt=dt)
    tol = 0.0 if exact else (1e-13 if T == np.float64 else 2e-5)
    for f, (_, h) in zip(Vo, V):
        assert_same(f, h, "V", rtol=tol)
    for c, (f, (_, h)) in enumerate(zip(tauo, tau)):
        assert_same(f, h, f"tau[{c}]", rtol=tol)
    assert_same(Pro, Pr[1], "Pr", rtol=tol)
    if wres:
        assert_same(dV[0], dV[1], "divV", rtol=tol)
        for D in range(nd):
            assert_same(rV[D][0], rV[D][1], f"r_V[{D}]", rtol=tol)
    # inputs are untouched
    for a, b in zip(inputs, before):
        assert torch.equal(a.data, b), f"input {a} was written"


@pytest.mark.parametrize("nd,T,exact", [(3, np.float64, True), (2, np.float64, True), (3, np.float32, True),
                                        (3, np.float64, False), (2, np.float64, False)])
def test_pt_solve_matches_oracle(arch, nd, T, exact):
    """40 PT iterations (fused kernel + free-slip pure-shear V batch) on a
    grid with a weak inclusion equal the oracle's update_stress! /
    update_velocity! / bc! sequence bitwise; the residual norms agree."""
    n = (22, 19, 17) if nd == 3 else (70, 45)
    g = chmy.UniformGrid(arch, origin=(-0.5,) * nd, extent=(1.0,) * nd, dims=n)
    og = ogrid(g)
    dtype = torch.float64 if T == np.float64 else torch.float32
    if nd == 3:
        eta = inclusion_setup_3d(g, radius=0.2)
    else:
        xs = [g.coords(Cc(), d, first=-1, count=g.dims[d] + 4) for d in range(2)]
        X, Y = xs[0][None, :], xs[1][:, None]
        eta = np.where(X ** 2 + Y ** 2 < 0.2 ** 2, 1e-2, 1.0)
    G, dt = 1.0, 0.25
    S = PTStokes(arch, g, dtype=dtype, eta_field=eta, G=G, dt=dt, normal_values=pure_shear_values(nd),
                 exact=exact)
    pure_shear_initial(g, S.V)
    S.apply_bc()
    S.sync_buffers()
    # oracle mirror
    tl = list(chmy.tensor_locations(nd).values())
    hV = [O.HField(og, tuple(l.code for l in f.loc), T) for f in S.V]
    for hv, f in zip(hV, S.V):
        hv.arr[...] = dev(f)
    htau = [O.HField(og, tuple(l.code for l in lc), T) for lc in tl]
    htold = [O.HField(og, tuple(l.code for l in lc), T) for lc in tl]
    hPr, hdV = O.HField(og, 0, T), O.HField(og, 0, T)
    hrV = [O.HField(og, tuple(l.code for l in f.loc), T) for f in S.V]
    heta, heml = O.HField(og, 0, T), O.HField(og, 0, T)
    heta.arr[...] = dev(S.eta)
    heml.arr[...] = dev(S.eta_ml)
    vals = pure_shear_values(nd)
    spec = []
    for D in range(nd):
        s = {d: O.Neumann() for d in range(nd)}
        s[D] = (O.Dirichlet(vals[D][0]), O.Dirichlet(vals[D][1]))
        spec.append((hV[D], s))
    p = O.pt_params(og)
    niter = 40
    for it in range(1, niter + 1):
        last = it == niter
        S.iteration(write_res=last)
        O.update_stress(og, htau, hPr, hdV, hV, htold, heta, 0.0, G, dt, p["dtau_Pr"], p["dtau_r"])
        O.update_velocity(og, hV, hrV, hPr, htau, None, 0.0, p["nudtau"], eta_ml=heml, G=G, dt=dt)
        O.bc_apply(og, spec)
    chmy.synchronize()
    tol = 0.0 if exact else 1e-12
    for f, h in zip(S.V, hV):
        assert_same(f, h, "V after 40 iterations", rtol=tol)
    for f, h in zip(S.tau, htau):
        assert_same(f, h, "tau after 40 iterations", rtol=tol)
    assert_same(S.Pr, hPr, "Pr after 40 iterations", rtol=tol)
    err = S.residual()
    O.bc_apply(og, [(hrV[D], {D: O.Dirichlet()}) for D in range(nd)])
    herr = [O.maxabs(hdV)] + [O.maxabs(h) for h in hrV]
    if exact:
        assert err == herr
    else:
        assert np.allclose(err, herr, rtol=1e-9, atol=1e-12)
