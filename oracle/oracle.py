"""ctypes front end of the CPU oracle (chmy_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  The oracle restates Chmy.jl's CPU semantics so
that the HIP library can be checked against it; only tests/,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this module.

Host fields here use Chmy's storage layout (src/Fields/field.jl:6-18,56-62):
``size = dims .+ 4H``, logical index ``I`` stored at ``I .+ 2H`` (1-based), x
fastest.  A numpy array of shape ``pad[::-1]`` in C order is exactly that.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libchmy_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class OField(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nd", C.c_int32), ("dtype", C.c_int32),
                ("n", C.c_int64 * 3), ("halo", C.c_int32 * 3), ("loc", C.c_int32 * 3)]


class OGrid(C.Structure):
    _fields_ = [("nd", C.c_int32), ("reserved", C.c_int32), ("n", C.c_int64 * 3),
                ("origin", C.c_double * 3), ("extent", C.c_double * 3),
                ("spacing", C.c_double * 3), ("inv_spacing", C.c_double * 3),
                ("axis", C.c_void_p * 3)]


AXIS_PAD = 4   # kPad in chmy_oracle.cpp


class OBox(C.Structure):
    _fields_ = [("lo", C.c_int64 * 3), ("len", C.c_int64 * 3)]


class OBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("vkind", C.c_int32), ("value", C.c_double),
                ("vfield", C.POINTER(OField))]


P = C.POINTER


def _declare(L):
    L.orc_last_error.argtypes = [C.c_char_p, C.c_int]
    L.orc_set_plant.argtypes = [C.c_int]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_plant.argtypes = []
    L.orc_vertex.restype = C.c_double
    L.orc_vertex.argtypes = [C.c_double, C.c_double, C.c_long]
    L.orc_center.restype = C.c_double
    L.orc_center.argtypes = [C.c_double, C.c_double, C.c_long]
    L.orc_coords.argtypes = [C.c_double, C.c_double, C.c_int, C.c_long, C.c_long, P(C.c_double)]
    L.orc_axis_coords.argtypes = [P(C.c_double), C.c_int, C.c_long, C.c_long, P(C.c_double)]
    L.orc_bc_side.argtypes = [P(OGrid), C.c_int, C.c_int, C.c_int, C.c_int, P(OField), P(OBC)]
    L.orc_bc_set.argtypes = [P(OGrid), C.c_int, P(C.c_int), P(OField), P(OBC)]
    L.orc_compute_q.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), C.c_double]
    L.orc_update_C.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), C.c_double]
    L.orc_update_old.argtypes = [P(OBox), C.c_int, P(OField), P(OField), P(OField), P(OField)]
    L.orc_update_stress.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), P(OField), P(OField), P(OField),
                                    P(OField)] + [C.c_double] * 6
    L.orc_update_velocity.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), P(OField), P(OField), P(OField),
                                      P(OField)] + [C.c_double] * 4 + [P(OField)]
    L.orc_stokes_residual.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), P(OField), P(OField), P(OField),
                                      P(OField), C.c_double, C.c_double, C.c_double, P(OField), P(OField),
                                      P(OField)]
    L.orc_stokes_grad_eta.argtypes = [P(OGrid), P(OField), P(OField), P(OField), P(OField), C.c_int]
    L.orc_thermal_flux.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), P(OField), C.c_double]
    L.orc_thermal.argtypes = [P(OGrid), P(OBox), P(OField), P(OField), P(OField), C.c_double]
    L.orc_maxloc.argtypes = [P(OBox), P(OField), P(OField)]
    L.orc_maxabs.restype = C.c_double
    L.orc_maxabs.argtypes = [P(OField), P(OBox)]
    L.orc_sumsq.restype = C.c_double
    L.orc_sumsq.argtypes = [P(OField), P(OBox)]
    L.orc_op.argtypes = [C.c_int, P(OGrid), P(OBox), P(OField), C.c_int, P(OField), P(OField), C.c_int]
    L.orc_op_masked.argtypes = [C.c_int, P(OGrid), P(OBox), P(OField), C.c_int, P(OField), P(OField), P(OField),
                                C.c_int]
    L.orc_sia_flux.argtypes = [P(OGrid), P(OField), P(OField), P(OField), P(OField), C.c_double, C.c_int,
                               C.c_double, C.c_double, C.c_double]
    L.orc_sia_update.argtypes = [P(OGrid), P(OField), P(OField), P(OField), P(OField), C.c_double, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_sia_Dmax.restype = C.c_double
    L.orc_sia_Dmax.argtypes = [P(OGrid), P(OField), P(OField), P(OField), C.c_double, C.c_int,
                               C.c_double, C.c_double, C.c_double]


def set_threads(n):
    """OpenMP threads of the oracle's loops (0: the count the process started with)."""
    _check(lib().orc_set_threads(int(n)))


def _check(rc):
    if rc:
        buf = C.create_string_buffer(256)
        lib().orc_last_error(buf, 256)
        raise RuntimeError(buf.value.decode())


CENTER, VERTEX = 0, 1

# Planted mistakes (chmy_oracle.cpp, enum Plant): used only by the tests that
# show each pin fails on a plausible error.
PLANTS = {
    "vertex_diff": 1, "lerp_knots": 2, "center_dirichlet": 3, "neumann_sign": 4, "bc_order": 5,
    "diff_flux": 6, "stress_dev": 7, "edge_swap": 8, "rv_swap": 9, "vel_ml_side": 10, "maxloc_2n": 11,
    "thermal_upwind": 12, "sia_exp": 13, "sia_grad": 14, "sia_m": 15, "resid_rp": 16, "pr_update": 17,
    "grad_edge": 18, "pt_params": 19, "vmag": 20, "divg": 21, "thermal_sign": 22, "mask_at_out": 23,
    "hlerp_arith": 24, "delta_sign": 25, "nu_diff_loc": 26, "nu_weight": 27, "nu_dirichlet": 28,
    "nu_neumann": 29,
}


class planted:
    """Context manager: run the oracle with one planted mistake."""

    def __init__(self, name):
        self.k = PLANTS[name]

    def __enter__(self):
        self.old = lib().orc_set_plant(self.k)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(self.old)
        return False


def _plant():
    return lib().orc_get_plant()


class Grid:
    """UniformGrid restated (src/Grids/uniform_axis.jl:7-12, structured_grid.jl:27-39).

    vertex_funcs[d] (optional) makes dim d a FunctionAxis
    (function_axis.jl:1-9): vertex(i) = f(i), evaluated here for
    i = -AXIS_PAD-1 .. n+2+AXIS_PAD; the oracle derives centres, spacings and
    weights from those values itself."""

    def __init__(self, origin, extent, dims, vertex_funcs=None):
        self.nd = len(dims)
        self.origin = tuple(float(o) for o in origin)
        self.extent = tuple(float(e) for e in extent)
        self.dims = tuple(int(d) for d in dims)
        self.spacing = tuple(e / n for e, n in zip(self.extent, self.dims))
        self.inv_spacing = tuple(1.0 / s for s in self.spacing)
        self.xv = [None] * self.nd
        for d, fn in enumerate(vertex_funcs or [None] * self.nd):
            if fn is not None:
                idx = range(-AXIS_PAD - 1, self.dims[d] + 3 + AXIS_PAD)
                self.xv[d] = np.array([float(fn(i)) for i in idx], dtype=np.float64)

    def size(self, loc):
        loc = _locs(loc, self.nd)
        return tuple(n + l for n, l in zip(self.dims, loc))

    def desc(self):
        g = OGrid()
        g.nd = self.nd
        for d in range(self.nd):
            g.n[d] = self.dims[d]
            g.origin[d] = self.origin[d]
            g.extent[d] = self.extent[d]
            g.spacing[d] = self.spacing[d]
            g.inv_spacing[d] = self.inv_spacing[d]
            if self.xv[d] is not None:
                g.axis[d] = self.xv[d].ctypes.data
        return g

    def coords(self, loc, dim, first=1, count=None):
        """coord(ax, loc, i) for i in first:first+count-1 (uniform_axis.jl:18-19;
        a FunctionAxis: abstract_axis.jl:18,20-21)."""
        if count is None:
            count = self.size(loc)[dim] if isinstance(loc, tuple) else self.dims[dim] + loc
        lc = loc[dim] if isinstance(loc, tuple) else loc
        out = np.empty(count, np.float64)
        if self.xv[dim] is not None:
            lib().orc_axis_coords(self.xv[dim].ctypes.data_as(P(C.c_double)), lc, first, count,
                                  out.ctypes.data_as(P(C.c_double)))
            return out
        lib().orc_coords(self.origin[dim], self.spacing[dim], lc, first, count,
                         out.ctypes.data_as(P(C.c_double)))
        return out

    def launcher_box(self):
        """Launcher iteration space [0, n+1] per dim (src/KernelLaunch.jl:41,109)."""
        return box([0] * self.nd, [n + 2 for n in self.dims])


def _locs(loc, nd):
    if isinstance(loc, int):
        return (loc,) * nd
    return tuple(loc)


def box(lo, length):
    b = OBox()
    for d in range(len(lo)):
        b.lo[d] = lo[d]
        b.len[d] = length[d]
    return b


class HField:
    """Host field in Chmy's padded layout."""

    def __init__(self, grid: Grid, loc, dtype=np.float64, halo=1):
        self.grid = grid
        self.nd = grid.nd
        self.loc = _locs(loc, grid.nd)
        self.halo = (halo,) * grid.nd if isinstance(halo, int) else tuple(halo)
        self.n = grid.size(self.loc)
        self.pad = tuple(n + 4 * h for n, h in zip(self.n, self.halo))
        self.arr = np.zeros(self.pad[::-1], dtype=dtype)
        self._d = None

    @property
    def dtype(self):
        return self.arr.dtype

    def desc(self):
        f = OField()
        f.data = self.arr.ctypes.data
        f.nd = self.nd
        f.dtype = 0 if self.arr.dtype == np.float64 else 1
        for d in range(self.nd):
            f.n[d] = self.n[d]
            f.halo[d] = self.halo[d]
            f.loc[d] = self.loc[d]
        return f

    def xyz(self):
        """Storage viewed with x first (Julia index order)."""
        return self.arr.T

    def interior(self, with_halo=False):
        sl = []
        for d in range(self.nd):
            h = self.halo[d]
            e = h if with_halo else 0
            sl.append(slice(2 * h - e, 2 * h + self.n[d] + e))
        return self.xyz()[tuple(sl)]

    def __getitem__(self, I):
        I = (I,) if isinstance(I, int) else I
        return self.xyz()[tuple(i + 2 * h - 1 for i, h in zip(I, self.halo))]

    def __setitem__(self, I, v):
        I = (I,) if isinstance(I, int) else I
        self.xyz()[tuple(i + 2 * h - 1 for i, h in zip(I, self.halo))] = v

    def set_interior(self, a):
        self.interior()[...] = a

    def copy(self):
        c = HField.__new__(HField)
        c.__dict__.update(self.__dict__)
        c.arr = self.arr.copy()
        return c


def _fields(fs):
    arr = (OField * max(1, len(fs)))()
    for i, f in enumerate(fs):
        arr[i] = f.desc()
    return arr


def _opt(f):
    return C.byref(f.desc()) if f is not None else None


# ------------------------------------------------------------ BCs
DIRICHLET, NEUMANN = 0, 1


class BC:
    def __init__(self, kind, value=None):
        self.kind = kind
        self.value = value

    def desc(self, keep):
        b = OBC()
        b.kind = self.kind
        if self.value is None:
            b.vkind = 0
        elif isinstance(self.value, HField):
            b.vkind = 2
            fd = self.value.desc()
            keep.append(fd)
            b.vfield = C.pointer(fd)
        else:
            b.vkind = 1
            b.value = float(self.value)
        return b


def Dirichlet(v=None):
    return BC(DIRICHLET, v)


def Neumann(v=None):
    return BC(NEUMANN, v)


def bc_side(grid, D, side, fields, bcs):
    keep = []
    fa = _fields(fields)
    ba = (OBC * max(1, len(bcs)))(*[b.desc(keep) for b in bcs])
    dt = 0 if fields[0].dtype == np.float64 else 1
    _check(lib().orc_bc_side(C.byref(grid.desc()), dt, D, side, len(fields), fa, ba))


def bc_apply(grid, pairs):
    """bc!(arch, grid, f => bc, ...) with per-field bc given as a BC (all
    dims/sides), a dict {dim: BC or (BC, BC)} (missing dims: nothing), per
    batch.jl:72-157.  Applies dims N..1, side 1 then 2, fields in order."""
    nd = grid.nd
    order = range(nd) if _plant() == PLANTS["bc_order"] else reversed(range(nd))
    for D in order:
        for s in (1, 2):
            fs, bs = [], []
            for f, spec in pairs:
                b = _side_bc(spec, D, s)
                if b is not None:
                    fs.append(f)
                    bs.append(b)
            if fs:
                bc_side(grid, D, s, fs, bs)


def _side_bc(spec, D, s):
    if spec is None:
        return None
    if isinstance(spec, BC):
        return spec
    b = spec.get(D)
    if b is None:
        return None
    if isinstance(b, tuple):
        return b[s - 1]
    return b


# ------------------------------------------------------------ kernels
def compute_q(grid, q, Cf, chi, b=None):
    b = b or grid.launcher_box()
    _check(lib().orc_compute_q(C.byref(grid.desc()), C.byref(b), _fields(q), C.byref(Cf.desc()), chi))


def update_C(grid, Cf, q, dt, b=None):
    b = b or grid.launcher_box()
    _check(lib().orc_update_C(C.byref(grid.desc()), C.byref(b), C.byref(Cf.desc()), _fields(q), dt))


def update_old(grid, T, tau, Told, tauo, b=None):
    b = b or grid.launcher_box()
    _check(lib().orc_update_old(C.byref(b), grid.nd, _opt(T), _fields(tau), _opt(Told), _fields(tauo)))


def update_stress(grid, tau, Pr, dV, V, tauo, eta, eta_ve, G, dt, dtau_Pr, dtau_r, b=None):
    """eta: scalar (reference form, eta_ve used as given) or HField (build extension)."""
    b = b or grid.launcher_box()
    ef = eta if isinstance(eta, HField) else None
    es = 0.0 if ef is not None else float(eta)
    _check(lib().orc_update_stress(C.byref(grid.desc()), C.byref(b), _fields(tau), C.byref(Pr.desc()),
                                   C.byref(dV.desc()), _fields(V), _fields(tauo), _opt(ef), es,
                                   float(eta_ve), G, dt, dtau_Pr, dtau_r))


def update_velocity(grid, V, rV, Pr, tau, rhog, eta_ve, nudtau, eta_ml=None, G=1.0, dt=1.0, b=None, fV=None):
    """fV: optional forcing added to r_V (build extension)."""
    b = b or grid.launcher_box()
    _check(lib().orc_update_velocity(C.byref(grid.desc()), C.byref(b), _fields(V), _fields(rV),
                                     C.byref(Pr.desc()), _fields(tau), _opt(rhog), _opt(eta_ml),
                                     float(eta_ve), G, dt, nudtau, _fields(fV) if fV is not None else None))


def stokes_residual(grid, V, tau, Pr, tauo, eta, rhog, G, dt, Rtau, RV, RP, b=None):
    """R_tau, R_V, R_P of a PT state (defect correction); eta scalar or HField."""
    b = b or grid.launcher_box()
    ef = eta if isinstance(eta, HField) else None
    es = 0.0 if ef is not None else float(eta)
    _check(lib().orc_stokes_residual(C.byref(grid.desc()), C.byref(b), _fields(V), _fields(tau), C.byref(Pr.desc()),
                                     _fields(tauo), _opt(ef), _opt(rhog), es, G, dt, _fields(Rtau), _fields(RV),
                                     C.byref(RP.desc())))


def stokes_grad_eta(grid, tau, tau_a, eta, grad, logscale=True):
    """dJ/d(log) eta at centres 0..n+1 from forward and adjoint-solve stresses."""
    _check(lib().orc_stokes_grad_eta(C.byref(grid.desc()), _fields(tau), _fields(tau_a), C.byref(eta.desc()),
                                     C.byref(grad.desc()), int(bool(logscale))))


def thermal_flux(grid, qT, T, V, lam, b=None):
    b = b or grid.launcher_box()
    _check(lib().orc_thermal_flux(C.byref(grid.desc()), C.byref(b), _fields(qT), C.byref(T.desc()),
                                  _fields(V), lam))


def thermal(grid, T, Told, qT, dt, b=None):
    b = b or grid.launcher_box()
    _check(lib().orc_thermal(C.byref(grid.desc()), C.byref(b), C.byref(T.desc()), C.byref(Told.desc()),
                             _fields(qT), dt))


def maxloc(grid, dst, src, b=None):
    b = b or grid.launcher_box()
    _check(lib().orc_maxloc(C.byref(b), C.byref(dst.desc()), C.byref(src.desc())))


def interior_box(f):
    return box([1] * f.nd, list(f.n))


def maxabs(f, b=None):
    b = b or interior_box(f)
    return lib().orc_maxabs(C.byref(f.desc()), C.byref(b))


def sumsq(f, b=None):
    b = b or interior_box(f)
    return lib().orc_sumsq(C.byref(f.desc()), C.byref(b))


OPS = {"partial": 0, "lerp": 1, "divg": 2, "vmag": 3, "d2": 4, "dkd": 5, "lapl": 6, "divg_grad": 7,
       "left": 8, "right": 9, "hlerp": 10, "delta": 11}


def op(name, grid, dst, src, k=None, dim=0, b=None):
    b = b or grid.launcher_box()
    src = src if isinstance(src, (list, tuple)) else [src]
    _check(lib().orc_op(OPS[name], C.byref(grid.desc()), C.byref(b), C.byref(dst.desc()), len(src),
                        _fields(src), _opt(k), dim))


def op_masked(name, grid, dst, src, masks, k=None, dim=0, b=None):
    """Masked operator (masked_operators.jl, masked_field_operators.jl);
    masks[q] = the FieldMask component at src[q]'s location."""
    b = b or grid.launcher_box()
    src = src if isinstance(src, (list, tuple)) else [src]
    masks = masks if isinstance(masks, (list, tuple)) else [masks]
    assert len(masks) == len(src)
    _check(lib().orc_op_masked(OPS[name], C.byref(grid.desc()), C.byref(b), C.byref(dst.desc()), len(src),
                               _fields(src), _fields(masks), _opt(k), dim))


# ------------------------------------------------------------ SIA
def sia_flux(grid, q, H, B, A=None, As=0.5, npow=3, mb=(0.5, 0.3, 0.25)):
    _check(lib().orc_sia_flux(C.byref(grid.desc()), _fields(q), C.byref(H.desc()), C.byref(B.desc()), _opt(A),
                              As, npow, *mb))


def sia_update(grid, H, Hout, q, B, dt, As=0.5, npow=3, mb=(0.5, 0.3, 0.25)):
    _check(lib().orc_sia_update(C.byref(grid.desc()), C.byref(H.desc()), C.byref(Hout.desc()), _fields(q),
                                C.byref(B.desc()), As, npow, *mb, dt))


def sia_Dmax(grid, H, B, A=None, As=0.5, npow=3, mb=(0.5, 0.3, 0.25)):
    return lib().orc_sia_Dmax(C.byref(grid.desc()), C.byref(H.desc()), C.byref(B.desc()), _opt(A), As, npow, *mb)


# ------------------------------------------------------------ PT parameters
def pt_params(grid, extent=None, re_m=2.3 * np.pi, r=0.5):
    """examples/stokes_3d_inc_ve_T.jl:105-112 (2D: stokes_2d_inc_ve_T.jl:87-94)."""
    L = extent or grid.extent
    lt = min(L) / re_m
    vdt = min(grid.spacing) / np.sqrt(grid.nd * 1.1)
    th = lt * (r + (1 / 3 if _plant() == PLANTS["pt_params"] else 4 / 3)) / vdt
    return dict(dtau_r=1.0 / (th + 1.0), nudtau=vdt * lt, dtau_Pr=r / th)
