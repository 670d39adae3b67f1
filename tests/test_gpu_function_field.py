"""set! and FunctionField evaluated on the device (chmy_set_program) against
the host evaluation of the same functions at the oracle's fma coordinates
(src/Fields/field.jl:119-142, src/Fields/function_field.jl:51-61).

Bar: bitwise where the function uses + - * / sqrt, comparisons, ifelse,
min/max, abs and floor (one IEEE operation each, in the same order on both
sides); where it calls exp, log, sin, cos, tanh or pow, the device math
library may differ from the host's by an ulp per call, so the bar is
|device - host| <= 8 eps max|host| (8 eps ~ 1.8e-15 relative)."""
import numpy as np
import pytest
import torch

import chmy
from chmy import Center as Cc
from chmy import Vertex as Vv
from chmy import _lib
from chmy.function_field import trace
import oracle as O
from _helpers import ogrid

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def init_incl2(x, y, x0, y0, r, inside, outside):
    """examples/stokes_2d_inc_ve_T.jl:106"""
    return chmy.ifelse((x - x0) ** 2 + (y - y0) ** 2 < r ** 2, inside, outside)


def init_incl3(x, y, z, x0, y0, z0, r, inside, outside):
    """examples/stokes_3d_inc_ve_T.jl:125"""
    return chmy.ifelse((x - x0) ** 2 + (y - y0) ** 2 + (z - z0) ** 2 < r ** 2, inside, outside)


def sia_bed(x, y):
    """examples/sia.jl:30-36"""
    bump1 = 0.3 * np.exp(-((x - 0.5) / 0.3) ** 2 - ((y - 0.0) / 0.5) ** 2)
    bump2 = 0.3 * np.exp(-((x + 0.5) / 0.3) ** 2 - ((y - 0.0) / 0.5) ** 2)
    bump3 = 1.0 * np.exp(-(x / 0.5) ** 2 - ((y - 0.5) / 0.5) ** 2)
    return bump1 + bump2 + bump3


def host_coords(og, loc, lo, length):
    """Coordinates (oracle's orc_coords, fma forms) of logical indices
    lo .. lo+length-1 along each dim, as an ij meshgrid."""
    axes = [og.coords(loc[d].code, d, first=lo[d], count=length[d]) for d in range(len(lo))]
    return np.meshgrid(*axes, indexing="ij")


def host_values(og, f, func, params, lo=None, length=None):
    lo = lo or (1,) * f.nd
    length = length or f.dims
    xs = host_coords(og, f.loc, lo, length)
    return np.broadcast_to(np.asarray(func(*xs, *params), dtype=np.float64), tuple(length))


def device_box(f, lo, length):
    sl = tuple(slice(lo[d] + 2 * f.halo[d] - 1, lo[d] + 2 * f.halo[d] - 1 + length[d]) for d in range(f.nd))
    return f.xyz()[sl].cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_set_continuous_inclusion_bitwise(arch, dtype):
    """The reference's set!(T, grid, init_incl; parameters=...) in 2D and 3D
    and its ρg at (Center, ..., Vertex): bitwise equal to the host
    evaluation at the same coordinates (then rounded to the field's dtype)."""
    for dims, func, loc, par in [
        ((62, 48), init_incl2, (Cc(), Vv()), dict(x0=0.0, y0=0.0, r=0.2, inside=1.0, outside=0.1)),
        ((62, 48), init_incl2, (Cc(), Cc()), (0.05, -0.1, 0.3, 2.5, -1.0)),
        ((30, 22, 17), init_incl3, (Cc(), Cc(), Vv()), dict(x0=0.0, y0=0.0, z0=0.0, r=0.2, inside=1.0, outside=0.0)),
    ]:
        nd = len(dims)
        g = chmy.UniformGrid(arch, origin=(-1.0,) * nd, extent=(2.0,) * nd, dims=dims)
        og = ogrid(g)
        f = chmy.Field(arch, g, loc, dtype)
        chmy.set_(f, func, g, parameters=par)
        want = host_values(og, f, func, tuple(par.values()) if isinstance(par, dict) else par)
        got = f.interior().cpu().numpy()
        assert np.array_equal(got, want.astype(got.dtype)), (dims, loc)
        # the halo is untouched: set! writes the interior only
        full = f.xyz().cpu().numpy().copy()
        full[tuple(slice(2 * h, 2 * h + m) for h, m in zip(f.halo, f.dims))] = 0
        assert not full.any()


def test_arithmetic_programs_bitwise(arch):
    """Every exact operation of the program against numpy on the host."""
    g = chmy.UniformGrid(arch, origin=(-1.5, -0.75), extent=(3.0, 2.0), dims=(70, 33))
    og = ogrid(g)
    # (device function, host form): Julia's literal powers x^3 = x*x*x and
    # x^-2 = inv(x)^2 are written out for numpy, whose ** calls pow
    funcs = [
        (lambda x, y: (x * y - 0.3) / (1.25 + y * y) + np.sqrt(np.abs(x)), None),
        (lambda x, y: np.minimum(x, y) - np.maximum(x * 0.5, -y) + np.floor(4.0 * x), None),
        (lambda x, y: chmy.ifelse((x > 0.1) & (y <= 0.2) | (x == y), x ** 3, y ** -1),
         lambda x, y: np.where((x > 0.1) & (y <= 0.2) | (x == y), x * x * x, 1.0 / y)),
        (lambda x, y: chmy.ifelse(~(x < y), -x, x ** 2 - y ** -2),
         lambda x, y: np.where(~(x < y), -x, x * x - (1.0 / y) * (1.0 / y))),
        (lambda x, y: 2.0 - x / 3.0 + (x != 0.25) * y, None),
    ]
    for k, (fn, hfn) in enumerate(funcs):
        f = chmy.Field(arch, g, (Vv(), Cc()))
        chmy.set_(f, fn, g)
        want = host_values(og, f, hfn or fn, ())
        got = f.interior().cpu().numpy()
        assert np.array_equal(got, want), k


def test_math_library_programs_within_ulps(arch):
    """exp, log, sin, cos, tanh, pow: within 8 eps of max|value| of numpy."""
    g = chmy.UniformGrid(arch, origin=(0.1, 0.2), extent=(2.0, 3.0), dims=(64, 40))
    og = ogrid(g)
    funcs = [sia_bed,
             lambda x, y: np.log(x + y) * np.tanh(x - y),
             lambda x, y: np.sin(3.0 * x) * np.cos(2.0 * y) + x ** 4.5,
             lambda x, y: np.power(x, y) - 2.0 ** x]
    for k, fn in enumerate(funcs):
        f = chmy.Field(arch, g, Cc())
        chmy.set_(f, fn, g)
        want = host_values(og, f, fn, ())
        got = f.interior().cpu().numpy()
        assert np.max(np.abs(got - want)) <= 8 * EPS * np.max(np.abs(want)), k


def test_function_field_materialized_over_its_storage(arch):
    """A FunctionField read in the halo (as kernels read ρg and the SIA bed)
    holds the function's values at every storage index, ghost and padding
    included; interior() is the copy-back path."""
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(40, 31))
    og = ogrid(g)
    B = chmy.FunctionField(sia_bed, g, Cc())
    f = B.materialize(arch)
    lo, ln = (-1, -1), (44, 35)
    want = host_values(og, f, sia_bed, (), lo, ln)
    got = device_box(f, lo, ln)
    assert np.max(np.abs(got - want)) <= 8 * EPS * np.max(np.abs(want))
    assert B.materialize(arch) is f          # coordinates only: evaluated once
    assert torch.equal(B.interior(arch), f.interior())
    # set!(f, other::AbstractField): the FunctionField's values index by index
    h = chmy.Field(arch, g, Cc())
    chmy.set_(h, B)
    assert torch.equal(h.interior(), f.interior())


def test_discrete_function_fields_with_field_parameters(arch):
    """sia.jl's discrete S = B + H (B a continuous FunctionField, H a Field)
    and ∂x S at x-vertices as a FunctionField of S: loads at index offsets and
    nested FunctionFields inlined at shifted indices; bitwise against the host
    composition from the materialized B (the loads and the subtraction are
    exact operations)."""
    n = (48, 37)
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=n)
    og = ogrid(g)
    rng = np.random.default_rng(3)
    H = chmy.Field(arch, g, Cc())
    hH = rng.random(H.data.shape)
    chmy.set_storage(H, hH)
    B = chmy.FunctionField(sia_bed, g, Cc())
    S = chmy.FunctionField(lambda g_, l, ix, iy, B, H: B[ix, iy] + H[ix, iy], g, Cc(), discrete=True,
                           parameters=(B, H))
    dSx = chmy.FunctionField(lambda g_, l, ix, iy, S: (S[ix, iy] - S[ix - 1, iy]) * chmy.inv_spacing(g_)[0],
                             g, (Vv(), Cc()), discrete=True, parameters=(S,))
    Bf = B.materialize(arch).xyz().cpu().numpy()
    Hs = hH.T
    Sf = Bf + Hs                                          # storage, (x, y)
    dx_inv = chmy.inv_spacing(g)[0]
    # vertex ix in 1..n+1 reads cells ix, ix-1 (storage ix+1, ix)
    want_dSx = (Sf[2:n[0] + 3, 2:-2] - Sf[1:n[0] + 2, 2:-2]) * dx_inv
    f = chmy.Field(arch, g, (Vv(), Cc()))
    chmy.set_(f, dSx)
    assert np.array_equal(f.interior().cpu().numpy(), want_dSx)
    # the discrete set! form with the same function
    f2 = chmy.Field(arch, g, (Vv(), Cc()))
    chmy.set_(f2, lambda g_, l, ix, iy, S: (S[ix, iy] - S[ix - 1, iy]) * chmy.inv_spacing(g_)[0], g,
              discrete=True, parameters=(S,))
    assert torch.equal(f2.data, f.data)
    # depends on H: re-evaluated at each use
    Sm = S.materialize(arch)
    assert np.array_equal(Sm.interior().cpu().numpy(), Sf[2:-2, 2:-2])
    H.data.add_(1.0)
    assert np.array_equal(S.materialize(arch).interior().cpu().numpy(), (Bf + (Hs + 1.0))[2:-2, 2:-2])
    # discrete index arithmetic: values of the indices themselves
    f3 = chmy.Field(arch, g, Cc())
    chmy.set_(f3, lambda g_, l, ix, iy: 1000.0 * ix + iy, g, discrete=True)
    I, J = np.meshgrid(np.arange(1, n[0] + 1), np.arange(1, n[1] + 1), indexing="ij")
    assert np.array_equal(f3.interior().cpu().numpy(), 1000.0 * I + J)


def test_rho_g_function_field_in_the_pt_solver(arch):
    """PTStokes with ρg given as the reference's FunctionField
    (examples/stokes_2d_inc_ve_T.jl:107) equals the solver with ρg set from
    the host-evaluated array, bitwise, over 20 iterations."""
    from chmy.stokes import PTStokes, pure_shear_initial, pure_shear_values
    n = (48, 40)
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=n)
    og = ogrid(g)
    par = dict(x0=0.0, y0=0.0, r=0.2, inside=1.0, outside=0.0)
    rg_ff = chmy.FunctionField(init_incl2, g, (Cc(), Vv()), parameters=par)
    rg_host = chmy.Field(arch, g, (Cc(), Vv()))
    lo, ln = (-1, -1), (n[0] + 4, n[1] + 5)
    rg_host.xyz()[...] = torch.from_numpy(host_values(og, rg_host, init_incl2, tuple(par.values()), lo, ln).copy())
    out = []
    for rg in (rg_ff, rg_host):
        S = PTStokes(arch, g, eta=1.0, G=1.0, dt=1.0, normal_values=pure_shear_values(2), rhog=rg)
        pure_shear_initial(g, S.V)
        S.apply_bc()
        S.sync_buffers()
        S.iterate(20)
        torch.cuda.synchronize()
        out.append([f.data.clone() for f in list(S.V) + list(S.tau) + [S.Pr]])
    for a, b in zip(*out):
        assert torch.equal(a, b)


def test_sia_bed_function_field(arch):
    """SIA with the bed as a FunctionField: the bed storage equals the
    function's device values and the forward run equals a run from that
    array, bitwise."""
    from chmy.sia import SIA
    n = (40, 36)
    g = chmy.UniformGrid(arch, origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=n)
    Bff = chmy.FunctionField(sia_bed, g, Cc())
    S1 = SIA(arch, g, Bff, npow=3, mb=(0.5, 0.3, 0.25))
    Barr = Bff.materialize(arch).data.cpu().numpy()
    S2 = SIA(arch, g, Barr, npow=3, mb=(0.5, 0.3, 0.25))
    assert torch.equal(S1.B.data, S2.B.data)
    for S in (S1, S2):
        S.reset()
        S.forward(30, [1e-4] * 30)
    assert torch.equal(S1.H.data, S2.H.data)
    assert float(S1.H.interior().abs().max()) > 0


def test_program_errors(arch):
    g = chmy.UniformGrid(arch, origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(16, 12))
    f = chmy.Field(arch, g, Cc())
    with pytest.raises(chmy.TraceError):
        chmy.set_(f, lambda x, y: x if x > y else y, g)
    with pytest.raises(chmy.TraceError):
        chmy.set_(f, lambda x, y: max(x, y), g)
    with pytest.raises(chmy.TraceError):
        def long_fn(x, y):
            v = x
            for k in range(200):
                v = v * 1.0001 + float(k)
            return v
        chmy.set_(f, long_fn, g)
    # a load past the source's storage is refused by the library before launch
    H = chmy.Field(arch, g, Cc())
    prog = trace(lambda g_, l, ix, iy, H: H[ix + 3, iy], g, Cc(), discrete=True, parameters=(H,))
    with pytest.raises(_lib.ChmyError, match="storage"):
        prog.run(g, f)
