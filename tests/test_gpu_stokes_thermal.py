"""The physical time loop of examples/stokes_3d_inc_ve_T.jl (two steps: the
second with the thermal update inside the PT loop) on the fused iteration
in exact form, against the oracle running the reference's kernel sequence
(update_old!, update_stress!, update_velocity! + V batch, update_thermal_flux!,
update_thermal! + T batch, residual check): same PT iteration counts and
bitwise equal V, Pr, tau, T."""
import math

import numpy as np
import pytest

import chmy
from chmy import Center as Cc, Vertex as Vv
from chmy.stokes_thermal import StokesThermal
import oracle as O
from _helpers import dev, ogrid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _incl(g, locs, inside, outside, r):
    """init_incl on the interior at the field's coordinates (stokes_3d_inc_ve_T.jl:125-127)."""
    nd = g.nd
    n = [g.dims[d] + (1 if locs[d] == "V" else 0) for d in range(nd)]
    xs = [g.coords(Vv() if locs[d] == "V" else Cc(), d, first=1, count=n[d]) for d in range(nd)]
    grids = np.meshgrid(*xs[::-1], indexing="ij")[::-1]   # storage order (z, y, x)
    r2 = sum(q * q for q in grids)
    arr = np.zeros(tuple(m + 4 for m in reversed(n)))
    arr[(slice(2, -2),) * nd] = np.where(r2 < r * r, inside, outside)
    return arr


@pytest.mark.parametrize("nd", [2, 3])
def test_time_loop_matches_reference_sequence(arch, nd):
    n = 12 if nd == 3 else 24
    L = 2.0
    g = chmy.UniformGrid(arch, origin=(-L / 2,) * nd, extent=(L,) * nd, dims=(n,) * nd)
    og = ogrid(g)
    eta, G = 10.0, 1.0
    lam = 1e-4 * L ** 2 / (eta / G)
    rg_loc = ("C",) * (nd - 1) + ("V",)
    rg_arr = _incl(g, rg_loc, 1.0, 0.0, 0.1 * L)
    T_arr = _incl(g, ("C",) * nd, 1.0, 0.1, 0.1 * L)
    rhog = chmy.Field(arch, g, tuple(Vv() if c == "V" else Cc() for c in rg_loc))
    chmy.set_storage(rhog, rg_arr)
    M = StokesThermal(arch, g, eta=eta, G=G, lam_rhoCp=lam, rhog=rhog, exact=True)
    chmy.set_storage(M.T, T_arr)
    M.apply_bc()
    res = M.run(nt=2, eps=1e-6)

    # oracle: the reference's kernel sequence
    C_, V_ = O.CENTER, O.VERTEX
    vl = [tuple(V_ if d == D else C_ for d in range(nd)) for D in range(nd)]
    tl = [(C_, C_), (C_, C_), (V_, V_)] if nd == 2 else \
        [(C_, C_, C_)] * 3 + [(V_, V_, C_), (V_, C_, V_), (C_, V_, V_)]
    Vf = [O.HField(og, l) for l in vl]
    rV = [O.HField(og, l) for l in vl]
    qT = [O.HField(og, l) for l in vl]
    tau = [O.HField(og, l) for l in tl]
    told = [O.HField(og, l) for l in tl]
    Pr, dV, T, Told = (O.HField(og, C_) for _ in range(4))
    rg = O.HField(og, tuple(V_ if c == "V" else C_ for c in rg_loc))
    rg.arr[...] = rg_arr
    T.arr[...] = T_arr
    p = O.pt_params(og)
    spec = []
    for D in range(nd):
        s = {d: O.Neumann() for d in range(nd)}
        s[D] = O.Dirichlet()
        spec.append((Vf[D], s))
    O.bc_apply(og, spec)
    O.bc_apply(og, [(T, O.Neumann())])
    iters = []
    ncheck, niter = 2 * n, 50 * n
    for it in (1, 2):
        O.update_old(og, T, tau, Told, told)
        d = og.spacing
        dt_diff = min(d) ** 2 / lam / nd / 2.1
        vmax = [O.maxabs(Vf[D]) for D in range(nd)]
        dt_adv = 0.1 * min(d[D] / vmax[D] if vmax[D] > 0 else math.inf for D in range(nd)) / nd / 2.1
        dt = min(dt_diff, dt_adv)
        eta_ve = 1.0 / (1.0 / eta + 1.0 / (G * dt))
        done = niter
        for q in range(1, niter + 1):
            O.update_stress(og, tau, Pr, dV, Vf, told, eta, eta_ve, G, dt, p["dtau_Pr"], p["dtau_r"])
            O.update_velocity(og, Vf, rV, Pr, tau, rg, eta_ve, p["nudtau"])
            O.bc_apply(og, spec)
            if it > 1:
                O.thermal_flux(og, qT, T, Vf, lam)
                O.thermal(og, T, Told, qT, dt)
                O.bc_apply(og, [(T, O.Neumann())])
            if q % ncheck == 0:
                O.bc_apply(og, [(rV[D], {D: O.Dirichlet()}) for D in range(nd)])
                err = [O.maxabs(dV) * eta / G] + [O.maxabs(rV[D]) * L / G for D in range(nd)]
                if all(e < 1e-6 for e in err):
                    done = q
                    break
        iters.append(done)
    assert [r[0] for r in res] == iters, (res, iters)
    # non-trivial: the buoyant inclusion drives a flow, both steps needed several
    # residual checks, and the second step moved T
    assert min(iters) >= 2 * ncheck, iters
    assert max(abs(float(O.maxabs(f))) for f in Vf) > 1e-5
    assert not np.array_equal(T.arr, T_arr)
    sl = (slice(1, -1),) * nd     # logical 0..n+2: interior, halos, boundary nodes
    for a, b in zip(list(M.S.V) + [M.S.Pr] + list(M.S.tau) + [M.T], Vf + [Pr] + tau + [T]):
        assert np.array_equal(dev(a)[sl], b.arr[sl])


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("nd", [2, 3])
def test_time_loop_pairs_equal_single_iterations(arch, exact, nd):
    """The thermal PT loop in passes of two iterations (the first
    iteration's V kept for its thermal update, chmy_stokes_iter2_first)
    against one iteration per launch: same iteration counts and residuals,
    bitwise equal V, Pr, tau and T.  An odd ncheck mixes passes with the
    single iterations before and at each check."""
    n = (21, 17, 14) if nd == 3 else (45, 38)
    L = 2.0
    g = chmy.UniformGrid(arch, origin=(-L / 2,) * nd, extent=(L,) * nd, dims=n)
    rg_loc = ("C",) * (nd - 1) + ("V",)
    rg_arr = _incl(g, rg_loc, 1.0, 0.0, 0.15 * L)
    T_arr = _incl(g, ("C",) * nd, 1.0, 0.1, 0.15 * L)
    runs = []
    for pairs in (True, False):
        rhog = chmy.Field(arch, g, tuple(Vv() if c == "V" else Cc() for c in rg_loc))
        chmy.set_storage(rhog, rg_arr)
        M = StokesThermal(arch, g, eta=10.0, G=1.0, rhog=rhog, exact=exact, pairs=pairs)
        chmy.set_storage(M.T, T_arr)
        M.apply_bc()
        res = M.run(nt=3, niter=160, ncheck=7, eps=1e-9)
        runs.append((res, [dev(f) for f in list(M.S.V) + [M.S.Pr] + list(M.S.tau) + [M.T]]))
    (ra, fa), (rb, fb) = runs
    assert ra == rb
    assert not np.array_equal(fa[-1], T_arr)
    for k, (x, y) in enumerate(zip(fa, fb)):
        assert np.array_equal(x.view(np.int64), y.view(np.int64)), k
