"""ctypes binding of libchmy_mi355x.so (include/chmy_mi355x.h).

The library is the product: every kernel call goes through it.  Loading
fails loudly when it is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CHMY_MI355X_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libchmy_mi355x.so")

F64, F32, BF16 = 0, 1, 2
CENTER, VERTEX = 0, 1
DIRICHLET, NEUMANN = 0, 1
VALUE_ZERO, VALUE_SCALAR, VALUE_FIELD = 0, 1, 2
OP_PARTIAL, OP_LERP, OP_DIVG, OP_VMAG, OP_LEFT, OP_RIGHT = 0, 1, 2, 3, 8, 9
OP_D2, OP_DKD, OP_LAPL, OP_DIVG_GRAD, OP_HLERP, OP_DELTA = 4, 5, 6, 7, 10, 11


class ChmyField(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nd", C.c_int32), ("dtype", C.c_int32),
                ("n", C.c_int64 * 3), ("halo", C.c_int32 * 3), ("loc", C.c_int32 * 3)]


class ChmyGrid(C.Structure):
    _fields_ = [("nd", C.c_int32), ("reserved", C.c_int32), ("n", C.c_int64 * 3),
                ("origin", C.c_double * 3), ("extent", C.c_double * 3),
                ("spacing", C.c_double * 3), ("inv_spacing", C.c_double * 3),
                ("axis", C.c_void_p * 3)]


# FunctionAxis metric tables (chmy_grid_t.axis)
AXIS_PAD, AXIS_ROWS = 4, 7
AX_VERTEX, AX_CENTER, AX_DC, AX_DV, AX_IDC, AX_IDV, AX_WCV = range(7)


class ChmyBox(C.Structure):
    _fields_ = [("lo", C.c_int64 * 3), ("len", C.c_int64 * 3)]


class ChmyPop(C.Structure):
    """One point-program operation (chmy_pop_t)."""
    _fields_ = [("op", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("c", C.c_int32)]


# point-program operation codes (chmy_set_program)
(P_CONST, P_INDEX, P_COORD, P_SPACING, P_INV_SPACING, P_LOAD) = 0, 1, 2, 3, 4, 5
(P_NEG, P_ABS, P_SQRT, P_EXP, P_LOG, P_SIN, P_COS, P_TANH, P_FLOOR, P_NOT) = range(8, 18)
(P_ADD, P_SUB, P_MUL, P_DIV, P_POW, P_MIN, P_MAX, P_LT, P_LE, P_GT, P_GE, P_EQ, P_NE, P_AND,
 P_OR) = range(24, 39)
P_POWI, P_SELECT, P_FMA = 40, 48, 49


class ChmyBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("vkind", C.c_int32), ("value", C.c_double),
                ("vfield", C.POINTER(ChmyField))]


P = C.POINTER
PF = P(ChmyField)
PG = P(ChmyGrid)
PB = P(ChmyBox)
VP = C.c_void_p
D = C.c_double

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "chmy_last_error": [],
    "chmy_init": [C.c_int],
    "chmy_device_count": [P(C.c_int)],
    "chmy_malloc_zeros": [P(C.c_void_p), C.c_size_t, VP],
    "chmy_free": [VP],
    "chmy_memset_zero": [VP, C.c_size_t, VP],
    "chmy_stream_create": [C.c_int, P(C.c_void_p)],
    "chmy_stream_destroy": [VP],
    "chmy_memcpy_h2d": [VP, VP, C.c_size_t, VP],
    "chmy_memcpy_d2h": [VP, VP, C.c_size_t, VP],
    "chmy_stream_sync": [VP],
    "chmy_coords": [PG, C.c_int, C.c_int, C.c_int64, C.c_int64, P(C.c_double)],
    "chmy_bc_side": [PG, C.c_int, C.c_int, C.c_int, PF, P(ChmyBC), VP],
    "chmy_bc_batchset": [PG, P(C.c_int32), PF, P(ChmyBC), VP],
    "chmy_op": [C.c_int, PG, PB, PF, C.c_int, PF, C.c_int, VP],
    "chmy_diffusion_flux": [PG, PB, PF, PF, D, VP],
    "chmy_diffusion_update": [PG, PB, PF, PF, D, VP],
    "chmy_diffusion_step": [PG, PB, PF, PF, PF, D, D, VP],
    "chmy_diffusion_iterate": [PG, PF, PF, PF, P(C.c_int32), P(ChmyBC), D, D, C.c_int, P(C.c_int), VP],
    "chmy_stokes_update_old": [PG, PB, PF, PF, PF, PF, VP],
    "chmy_stokes_stress": [PG, PB, PF, PF, PF, PF, PF, PF, D, D, D, D, D, D, VP],
    "chmy_stokes_velocity": [PG, PB, PF, PF, PF, PF, PF, PF, D, D, D, D, VP],
    "chmy_maxloc": [PG, PB, PF, PF, VP],
    "chmy_thermal_flux": [PG, PB, PF, PF, PF, D, VP],
    "chmy_thermal_update": [PG, PB, PF, PF, PF, D, VP],
    "chmy_thermal_step": [PG, PB, PF, PF, PF, PF, PF, D, D, VP],
    "chmy_stokes_iter": [PG, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, C.c_int, D, D, D, D, D, D, D,
                         VP],
    "chmy_stokes_iterate": [PG, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, P(C.c_int32),
                            P(C.c_int32), P(ChmyBC), C.c_int, C.c_int, C.c_int, D, D, D, D, D, D, D, P(C.c_int),
                            VP],
    "chmy_stokes_iter2": [PG, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, P(C.c_int32), P(C.c_int32),
                          P(ChmyBC), C.c_int, D, D, D, D, D, D, D, VP],
    "chmy_stokes_iter2_first": [PG, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, PF, P(C.c_int32), P(C.c_int32),
                                P(ChmyBC), C.c_int, D, D, D, D, D, D, D, VP],
    "chmy_stokes_residual": [PG, PF, PF, PF, PF, PF, PF, PF, D, D, D, PF, PF, PF, PF, PF, VP],
    "chmy_stokes_accumulate": [PG, PF, PF, PF, PF, PF, PF, VP],
    "chmy_stokes_grad_eta": [PG, PF, PF, PF, PF, C.c_int, VP],
    "chmy_sia_flux": [PG, PF, PF, PF, PF, D, C.c_int, VP],
    "chmy_sia_update": [PG, PF, PF, PF, PF, D, D, D, D, VP],
    "chmy_sia_diffusivity": [PG, PF, PF, PF, PF, D, C.c_int, VP],
    "chmy_sia_step": [PG, PF, PF, PF, PF, D, C.c_int, D, D, D, D, VP],
    "chmy_sia_step_adjoint": [PG, PF, PF, PF, PF, PF, PF, D, C.c_int, D, D, D, D, PF, VP],
    "chmy_sia_run": [PG, PF, C.c_int, PF, PF, D, C.c_int, D, D, D, P(C.c_double), VP],
    "chmy_sia_reverse": [PG, PF, C.c_int, C.c_int, C.c_int, PF, PF, PF, PF, PF, PF, D, C.c_int, D, D, D,
                         P(C.c_double), P(C.c_int), VP],
    "chmy_sia_reverse_noalloc": [PG, PF, C.c_int, C.c_int, C.c_int, PF, PF, PF, PF, D, C.c_int, D, D, D,
                                 P(C.c_double), VP],
    "chmy_op_masked": [C.c_int, PG, PB, PF, C.c_int, PF, PF, PF, C.c_int, VP],
    "chmy_set_program": [PG, PF, PB, P(ChmyPop), C.c_int, P(C.c_double), C.c_int, PF, C.c_int, VP],
    "chmy_maxabs_dev": [PF, PB, P(C.c_double), VP],
    "chmy_maxabs": [PF, PB, P(C.c_double), VP],
    "chmy_sumsq": [PF, PB, P(C.c_double), VP],
}
_RESTYPE = {"chmy_last_error": C.c_char_p}

_lib = None


class ChmyError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChmyError(f"{LIB_PATH} not found: build it with `make -C chmy.jl_amd` "
                            "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def call(name, *args):
    """Call an entry point and raise ChmyError with chmy_last_error() on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.chmy_last_error().decode(errors="replace")
        raise ChmyError(f"{name} failed ({rc}): {msg}")
    return rc
