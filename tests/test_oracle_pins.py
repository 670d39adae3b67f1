"""More pins of the CPU oracle: closed forms and brute force for the parts
that the reference's unit tests do not reach.

- SIA diffusivity (examples/sia.jl:46-49, with the per-cell rate factor A of
  the build): a planar bed under uniform ice gives q = -D s with
  D = (A h^(n+2) + As h^n) |s|^(n-1) at every interior face.
- Variable-viscosity stress update (build extension of
  examples/stokes_3d_inc_ve_T.jl:23-46): with tau = tau_old = 0 and an affine
  V, one update_stress! step stores 2 eps' eta_ve(eta_loc) dtau_r, so eta_loc
  is recovered per component and must equal an affine eta at that
  component's own coordinates (Center for xx/yy/zz, the xy/xz/yz edges).
- Velocity pseudo-step coefficient (build extension): eta_ve of
  max(eta_ml[I-e_D], eta_ml[I]) equals eta_ve of the maximum of eta over the
  union of the two cells' 3^N neighbourhoods, computed by brute force.
- maxloc against a brute-force 3^N maximum.
- update_stress! on an affine V with nonzero divergence and uniform tau,
  tau_old, Pr (examples/stokes_3d_inc_ve_T.jl:32-45): divV, Pr and every
  tau component in closed form, including the deviatoric -divV/3 term.
- update_velocity! on affine Pr and stresses (examples/stokes_3d_inc_ve_T.jl:
  51-56): r_V is exact, which fixes which stress component and which
  derivative enters each velocity component.
- bc! order with two Dirichlet values (batch.jl:20-29): the corner ghost of
  dims y then x differs from x then y.
- PT parameters (examples/stokes_3d_inc_ve_T.jl:105-112): with them the
  homogeneous pure-shear problem converges at the rate the damped-wave PT
  method gives; a wrong theta converges markedly slower.
tests/test_oracle_planted.py shows each of these fails on a planted mistake.
"""
import numpy as np
import pytest

import oracle as O
from oracle import CENTER as C, VERTEX as V


def vloc(D, nd):
    return tuple(V if d == D else C for d in range(nd))


def _coords_full(g, f):
    """Coordinates of every stored point of f per dim (logical -1 .. n+2)."""
    return [g.coords(f.loc, d, first=1 - 2 * f.halo[d], count=f.n[d] + 4 * f.halo[d]) for d in range(g.nd)]


def _set_full(g, f, fun):
    """f = fun(x, y[, z]) on the whole padded storage."""
    xs = _coords_full(g, f)
    grids = np.meshgrid(*xs, indexing="ij")
    f.xyz()[...] = fun(*grids)


# ---------------------------------------------------------------- SIA D
@pytest.mark.parametrize("npow", [2, 3, 4])
@pytest.mark.parametrize("slope", [(0.3, 0.0), (0.0, -0.2), (0.12, 0.16)])
def test_sia_flux_planar_bed_power_law(npow, slope):
    """q_x = -D s_x, q_y = -D s_y on a planar bed B = s·x under uniform
    H = h and A, with D = (A h^(n+2) + As h^n) |s|^(n-1); the powers are
    compared with numpy's correctly rounded pow (the oracle multiplies
    repeatedly, Julia's Float64^Int is compensated: both are within a few ulp
    of the exact power)."""
    n = 10
    g = O.Grid((0.0, 0.0), (1.0, 1.0), (n, n))
    h, Aval, As = 0.7, 2.5, 0.4
    sx, sy = slope
    H, B, A = O.HField(g, C), O.HField(g, C), O.HField(g, C)
    H.arr[...] = h
    A.arr[...] = Aval
    _set_full(g, B, lambda x, y: sx * x + sy * y)
    q = [O.HField(g, vloc(D, 2)) for D in range(2)]
    O.sia_flux(g, q, H, B, A, As=As, npow=npow)
    smag = np.hypot(sx, sy)
    D = (Aval * h ** (npow + 2) + As * h ** npow) * smag ** (npow - 1)
    for d, s in enumerate((sx, sy)):
        got = q[d].interior()[:n, :n]
        assert np.allclose(got, -D * s, rtol=1e-12, atol=1e-15 * abs(D)), (d, got.flat[0], -D * s)


# ---------------------------------------------------------------- variable viscosity
ETA = (1.5, 0.3, 0.2, 0.1)   # eta = a + b x + c y + d z, positive on [-0.5, 0.5]^3


def _eta_affine(*X):
    v = ETA[0]
    for k, x in enumerate(X):
        v = v + ETA[k + 1] * x
    return v


def _tau_locs(nd):
    return [(C, C), (C, C), (V, V)] if nd == 2 else [(C, C, C)] * 3 + [(V, V, C), (V, C, V), (C, V, V)]


@pytest.mark.parametrize("nd", [2, 3])
def test_stress_viscosity_locations_affine_eta(nd):
    """eta at each stress component's location is lerp(eta) to that location:
    for an affine eta it is eta evaluated at the component's coordinates."""
    dims = (7, 6, 5)[:nd]
    g = O.Grid((-0.5,) * nd, (1.0,) * nd, dims)
    al, be, ga, de = 0.8, -0.6, 0.5, 0.4        # Vx = al y + be z + de x, Vy = ga z
    Vf = [O.HField(g, vloc(D, nd)) for D in range(nd)]
    if nd == 2:
        _set_full(g, Vf[0], lambda x, y: al * y + de * x)
    else:
        _set_full(g, Vf[0], lambda x, y, z: al * y + be * z + de * x)
        _set_full(g, Vf[1], lambda x, y, z: ga * z)
    eta = O.HField(g, C)
    _set_full(g, eta, _eta_affine)
    tl = _tau_locs(nd)
    tau = [O.HField(g, l) for l in tl]
    tauo = [O.HField(g, l) for l in tl]
    Pr, dV = O.HField(g, C), O.HField(g, C)
    G, dt = 1.0, 0.8
    Gdt = G * dt
    O.update_stress(g, tau, Pr, dV, Vf, tauo, eta, 0.0, G, dt, 0.0, 1.0)
    # rate of each component: tau = 2 eps' eta_ve (dtau_r = 1)
    if nd == 2:
        rate = {0: 2 * (de - de / 3), 1: 2 * (0 - de / 3), 2: al}
    else:
        rate = {0: 2 * (de - de / 3), 1: 2 * (0 - de / 3), 2: 2 * (0 - de / 3), 3: al, 4: be, 5: ga}
    for c, l in enumerate(tl):
        f = tau[c]
        got = f.interior()
        eta_ve = got / rate[c]
        eta_loc = 1.0 / (1.0 / eta_ve - 1.0 / Gdt)
        xs = [g.coords(l, d) for d in range(nd)]
        want = _eta_affine(*np.meshgrid(*xs, indexing="ij"))
        assert np.allclose(eta_loc, want, rtol=1e-11, atol=0), (c, np.max(np.abs(eta_loc - want)))


def _brute_union_max(eta, D, I, nd):
    """max of eta over cells I-e_D-1 .. I+1 in dim D and I-1 .. I+1 in the
    others (the union of the 3^N neighbourhoods of cells I-e_D and I)."""
    sl = []
    for d in range(nd):
        lo = I[d] - (2 if d == D else 1)
        hi = I[d] + 1
        sl.append(slice(lo + 1, hi + 2))        # storage index = logical + 1 (halo 1)
    return np.max(eta.xyz()[tuple(sl)])


@pytest.mark.parametrize("nd", [2, 3])
def test_velocity_coefficient_union_max(nd):
    """V += r_V nudtau / eta_ve(max(eta_ml[I-e_D], eta_ml[I])): with tau = 0
    and an affine Pr, r_V = -grad Pr exactly, so each velocity point yields
    its eta, which must be the brute-force maximum of eta over the union of
    its two cells' neighbourhoods."""
    dims = (6, 5, 4)[:nd]
    g = O.Grid((-0.5,) * nd, (1.0,) * nd, dims)
    rng = np.random.default_rng(5 + nd)
    eta = O.HField(g, C)
    eta.arr[...] = rng.uniform(0.5, 3.0, eta.arr.shape)
    eml = O.HField(g, C)
    O.maxloc(g, eml, eta)
    p = (0.7, -1.1, 0.9)[:nd]
    Pr = O.HField(g, C)
    _set_full(g, Pr, lambda *X: sum(pk * x for pk, x in zip(p, X)))
    Vf = [O.HField(g, vloc(D, nd)) for D in range(nd)]
    rV = [O.HField(g, vloc(D, nd)) for D in range(nd)]
    tau = [O.HField(g, l) for l in _tau_locs(nd)]
    G, dt, nu = 1.0, 0.6, 0.3
    O.update_velocity(g, Vf, rV, Pr, tau, None, 0.0, nu, eta_ml=eml, G=G, dt=dt)
    for D in range(nd):
        got = Vf[D].interior()
        eta_ve = -p[D] * nu / got
        eta_v = 1.0 / (1.0 / eta_ve - 1.0 / (G * dt))
        want = np.empty_like(eta_v)
        for I in np.ndindex(*want.shape):
            want[I] = _brute_union_max(eta, D, tuple(i + 1 for i in I), nd)
        assert np.allclose(eta_v, want, rtol=1e-11, atol=0), D


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_maxloc_brute_force(nd):
    """eta_ml[I] = max of eta over I-1..I+1 in every dim, over the Launcher
    box [0, n+1]^N (exact: max does not round)."""
    dims = (7, 6, 5)[:nd]
    g = O.Grid((0.0,) * nd, (1.0,) * nd, dims)
    rng = np.random.default_rng(17)
    src = O.HField(g, C)
    src.arr[...] = rng.standard_normal(src.arr.shape)
    dst = O.HField(g, C)
    O.maxloc(g, dst, src)
    S = src.xyz()
    got = dst.xyz()
    for I in np.ndindex(*tuple(n + 2 for n in dims)):
        st = tuple(i + 1 for i in I)             # logical I -> storage I + 1
        want = np.max(S[tuple(slice(s - 1, s + 2) for s in st)])
        assert got[st] == want, I


# ---------------------------------------------------------------- update_stress! / update_velocity!
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("visc", ["scalar", "field"])
def test_stress_update_affine_velocity_closed_form(nd, visc):
    """V_D = a_D x_D plus off-diagonal shears, uniform tau, tau_old and Pr:
    divV = sum a_D; Pr' = Pr - divV eta_ve dtau_Pr;
    tau_DD' = tau + (-(tau - tau_old)/(G dt) - tau/eta + 2 (a_D - divV/3)) eta_ve dtau_r;
    tau_DE' = tau + (-(tau - tau_old)/(G dt) - tau/eta + 2 eps_DE) eta_ve dtau_r."""
    dims = (6, 5, 4)[:nd]
    g = O.Grid((-0.5,) * nd, (1.0,) * nd, dims)
    a = (0.9, -0.4, 0.7)[:nd]
    sh = {(0, 1): 0.3, (1, 0): -0.5, (0, 2): 0.25, (2, 0): 0.6, (1, 2): -0.35, (2, 1): 0.45}
    Vf = [O.HField(g, vloc(D, nd)) for D in range(nd)]
    for D in range(nd):
        def fun(*X, D=D):
            v = a[D] * X[D]
            for E in range(nd):
                if E != D:
                    v = v + sh[(D, E)] * X[E]
            return v
        _set_full(g, Vf[D], fun)
    tl = _tau_locs(nd)
    t0 = [0.2, -0.3, 0.15, 0.4, -0.25, 0.1][:len(tl)]
    to = [-0.1, 0.05, 0.3, -0.2, 0.35, -0.15][:len(tl)]
    tau = [O.HField(g, l) for l in tl]
    tauo = [O.HField(g, l) for l in tl]
    for c in range(len(tl)):
        tau[c].arr[...] = t0[c]
        tauo[c].arr[...] = to[c]
    Pr, dV = O.HField(g, C), O.HField(g, C)
    p0 = 0.37
    Pr.arr[...] = p0
    eta_s, G, dt = 0.8, 1.0, 0.45
    eve = 1.0 / (1.0 / eta_s + 1.0 / (G * dt))
    pp = O.pt_params(g)
    eta = eta_s
    if visc == "field":
        eta = O.HField(g, C)
        eta.arr[...] = eta_s
    O.update_stress(g, tau, Pr, dV, Vf, tauo, eta, eve, G, dt, pp["dtau_Pr"], pp["dtau_r"])
    dv = sum(a)
    assert np.allclose(dV.interior(), dv, rtol=1e-13, atol=1e-14)
    assert np.allclose(Pr.interior(), p0 - dv * eve * pp["dtau_Pr"], rtol=1e-13, atol=1e-14)
    off = [(0, 1)] if nd == 2 else [(0, 1), (0, 2), (1, 2)]
    for c in range(len(tl)):
        if c < nd:
            eps = 2 * (a[c] - dv / 3)
        else:
            D, E = off[c - nd]
            eps = 2 * 0.5 * (sh[(D, E)] + sh[(E, D)])
        r = -(t0[c] - to[c]) / (G * dt) - t0[c] / eta_s + eps
        want = t0[c] + r * eve * pp["dtau_r"]
        assert np.allclose(tau[c].interior(), want, rtol=1e-12, atol=1e-14), (c, tau[c].interior().flat[0], want)


@pytest.mark.parametrize("nd", [2, 3])
def test_velocity_residual_affine_stresses(nd):
    """r_V = -grad Pr + div tau - rho g on affine Pr and stresses, each
    stress component with its own slope per dim: r_Vx = -p_x + s_xx,x +
    s_xy,y (+ s_xz,z), and so on."""
    dims = (6, 5, 4)[:nd]
    g = O.Grid((-0.5,) * nd, (1.0,) * nd, dims)
    rng = np.random.default_rng(23 + nd)
    tl = _tau_locs(nd)
    S = rng.uniform(-1, 1, (len(tl), nd))
    p = rng.uniform(-1, 1, nd)
    Pr = O.HField(g, C)
    _set_full(g, Pr, lambda *X: sum(pk * x for pk, x in zip(p, X)))
    tau = [O.HField(g, l) for l in tl]
    for c in range(len(tl)):
        _set_full(g, tau[c], lambda *X, c=c: sum(S[c, d] * X[d] for d in range(nd)))
    rg = O.HField(g, vloc(nd - 1, nd))
    r0 = 0.6
    rg.arr[...] = r0
    Vf = [O.HField(g, vloc(D, nd)) for D in range(nd)]
    rV = [O.HField(g, vloc(D, nd)) for D in range(nd)]
    O.update_velocity(g, Vf, rV, Pr, tau, rg, 1.0, 1.0)
    if nd == 2:   # tau = (xx, yy, xy)
        want = [-p[0] + S[0, 0] + S[2, 1], -p[1] + S[1, 1] + S[2, 0] - r0]
    else:         # tau = (xx, yy, zz, xy, xz, yz)
        want = [-p[0] + S[0, 0] + S[3, 1] + S[4, 2],
                -p[1] + S[1, 1] + S[3, 0] + S[5, 2],
                -p[2] + S[2, 2] + S[4, 0] + S[5, 1] - r0]
    for D in range(nd):
        assert np.allclose(rV[D].interior(), want[D], rtol=1e-12, atol=1e-13), D
        assert np.allclose(Vf[D].interior(), want[D], rtol=1e-12, atol=1e-13), D


# ---------------------------------------------------------------- bc! order
def test_bc_order_two_dirichlet_corner():
    """Dirichlet(5) in x and Dirichlet(3) in y on a Center field: y runs
    first (dims N..1), so the corner ghost is 2·5 - (2·3 - f[1,1]) = 4 + f[1,1]
    (x first would give f[1,1] - 4)."""
    g = O.Grid((0, 0), (1, 1), (4, 4))
    f = O.HField(g, C)
    f.set_interior(np.arange(16, dtype=float).reshape(4, 4, order="F") + 1)
    O.bc_apply(g, [(f, {0: O.Dirichlet(5.0), 1: O.Dirichlet(3.0)})])
    assert f[0, 0] == 4.0 + f[1, 1]
    assert f[5, 5] == 4.0 + f[4, 4]
    assert f[1, 0] == 6.0 - f[1, 1] and f[0, 1] == 10.0 - f[1, 1]


# ---------------------------------------------------------------- PT parameters
def _pt_pure_shear_iterations(n, eps=1e-8, cap=6000):
    g = O.Grid((-0.5, -0.5), (1.0, 1.0), (n, n))
    Vf = [O.HField(g, vloc(D, 2)) for D in range(2)]
    rV = [O.HField(g, vloc(D, 2)) for D in range(2)]
    tl = _tau_locs(2)
    tau = [O.HField(g, l) for l in tl]
    tauo = [O.HField(g, l) for l in tl]
    Pr, dV = O.HField(g, C), O.HField(g, C)
    spec = [(Vf[0], {0: (O.Dirichlet(1.0), O.Dirichlet(-1.0)), 1: O.Neumann()}),
            (Vf[1], {0: O.Neumann(), 1: (O.Dirichlet(-1.0), O.Dirichlet(1.0))})]
    O.bc_apply(g, spec)
    p = O.pt_params(g)
    G, dt, eta = 1.0, 1e3, 1.0
    eve = 1.0 / (1.0 / eta + 1.0 / (G * dt))
    for it in range(1, cap + 1):
        O.update_stress(g, tau, Pr, dV, Vf, tauo, eta, eve, G, dt, p["dtau_Pr"], p["dtau_r"])
        O.update_velocity(g, Vf, rV, Pr, tau, None, eve, p["nudtau"])
        O.bc_apply(g, spec)
        if it % n == 0:
            O.bc_apply(g, [(rV[0], {0: O.Dirichlet()}), (rV[1], {1: O.Dirichlet()})])
            err = max([O.maxabs(dV)] + [O.maxabs(r) for r in rV])
            if not np.isfinite(err):
                return cap + 1
            if err < eps:
                return it
    return cap + 1


def test_pt_params_converge_at_the_damped_wave_rate():
    """With the reference's PT parameters (re_m = 2.3π, r = 0.5) the PT
    iteration of the homogeneous pure-shear problem from V = 0 reaches 1e-8
    within 30 n iterations at n = 32 (25 n measured at n = 16, 32, 48: the
    damped-wave method converges in O(n) iterations at this Reynolds number;
    theta with r + 1/3 in place of r + 4/3 takes over 200 n)."""
    n = 32
    it = _pt_pure_shear_iterations(n)
    assert it <= 30 * n, it


# ---------------------------------------------------------------- δ and hlerp
def _rand_field(g, loc, rng, lo, hi):
    f = O.HField(g, loc)
    f.arr[...] = rng.uniform(lo, hi, f.arr.shape)
    return f


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_delta_is_the_partial_before_its_inverse_spacing(nd):
    """δ (partial_derivatives.jl:2): ∂ = δ·iΔ (:5), so δ times iΔ is the
    reference-pinned ∂ bitwise, for every source location and dim."""
    g = O.Grid((-0.3,) * nd, (1.3,) * nd, [(11,), (9, 7), (7, 6, 5)][nd - 1])
    rng = np.random.default_rng(nd)
    for m in range(1 << nd):
        loc = tuple((m >> d) & 1 for d in range(nd))
        f = _rand_field(g, loc, rng, -1.0, 1.0)
        for D in range(nd):
            to = tuple(1 - l if d == D else l for d, l in enumerate(loc))
            a, b = O.HField(g, to), O.HField(g, to)
            O.op("delta", g, a, f, dim=D)
            O.op("partial", g, b, f, dim=D)
            assert np.array_equal(a.arr * g.inv_spacing[D], b.arr), (loc, D)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_hlerp_is_the_harmonic_rule(nd):
    """hlerp (interpolation.jl:15,101): each knot pair a, b reduces to
    inv(muladd(0.5, inv(b) - inv(a), inv(a))), the harmonic mean 2ab/(a+b).
    In 1D hlerp(f) is 1/lerp(1/f) bitwise, lerp being reference-pinned; in
    2D/3D the nested harmonic means equal the harmonic mean of the 2^m knots,
    1/mean(1/knots), to a few ulp."""
    g = O.Grid((-0.3,) * nd, (1.3,) * nd, [(11,), (9, 7), (7, 6, 5)][nd - 1])
    rng = np.random.default_rng(10 + nd)
    for m in range(1 << nd):
        loc = tuple((m >> d) & 1 for d in range(nd))
        f = _rand_field(g, loc, rng, 0.2, 5.0)
        inv = O.HField(g, loc)
        inv.arr[...] = 1.0 / f.arr
        for q in range(1 << nd):
            to = tuple((q >> d) & 1 for d in range(nd))
            h, l = O.HField(g, to), O.HField(g, to)
            box = O.interior_box(h)
            O.op("hlerp", g, h, f, b=box)
            O.op("lerp", g, l, inv, b=box)
            sl = tuple(slice(2, 2 + n) for n in reversed(h.n))
            want = 1.0 / l.arr[sl]
            if nd == 1:
                assert np.array_equal(h.arr[sl], want), (loc, to)
            else:
                assert np.max(np.abs(h.arr[sl] / want - 1.0)) <= 8e-16 * (1 << nd), (loc, to)
