"""Host side of set!/FunctionField on the device: the tracer's point programs
(operations, deduplication, Julia's literal powers, nested FunctionFields at
shifted indices) and the library's validation of programs before any launch.
No GPU: programs are checked, not run."""
import ctypes as C
import re
import os

import numpy as np
import pytest

import chmy
from chmy import Center as Cc
from chmy import Vertex as Vv
from chmy import _lib
from chmy.function_field import P, TraceError, trace

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_opcodes_match_the_header():
    src = open(os.path.join(ROOT, "include", "chmy_mi355x.h")).read()
    enum = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"CHMY_P_(\w+)\s*=\s*(\d+)", src))
    assert enum and set(enum) == set(P)
    for k, v in P.items():
        assert enum[k] == v, k


def test_inclusion_program():
    """init_incl (examples/stokes_2d_inc_ve_T.jl:106) at (Center, Vertex):
    two coordinates at their own locations, literal squares as products,
    one comparison and one select, constants deduplicated."""
    g = chmy.UniformGrid(origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(8, 6))
    f = lambda x, y, x0, y0, r, i, o: chmy.ifelse((x - x0) ** 2 + (y - y0) ** 2 < r ** 2, i, o)
    p = trace(f, g, (Cc(), Vv()), parameters=dict(x0=0.0, y0=0.0, r=0.2, i=1.0, o=0.0))
    ops = [o[0] for o in p.ops]
    assert p.ops[0] == (P["COORD"], 0, 0, 0) and p.ops[1] == (P["COORD"], 1, 1, 0)
    assert ops.count(P["POWI"]) == 2 and P["POW"] not in ops and P["MUL"] not in ops
    assert ops[-1] == P["SELECT"] and ops.count(P["LT"]) == 1
    # x0 = y0 = o = 0.0 share one constant; r**2 is folded on the host as r*r
    assert sorted(p.consts) == [0.0, 0.2 * 0.2, 1.0]


def test_repeated_subexpressions_are_computed_once():
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(4, 4))
    p = trace(lambda x, y: np.exp(-(x * y) ** 2) + np.exp(-(x * y) ** 2), g, Cc())
    ops = [o[0] for o in p.ops]
    assert ops.count(P["EXP"]) == 1 and ops.count(P["MUL"]) == 1


def test_powers_follow_julia_literal_pow():
    g = chmy.UniformGrid(origin=(0.0,), extent=(1.0,), dims=(4,))
    for e, want in [(2, ("POWI", 2)), (3, ("POWI", 3)), (-1, ("POWI", -1)), (-2, ("POWI", -2)), (4, ("POW",)),
                    (0.5, ("POW",)), (2.0, ("POW",))]:
        p = trace(lambda x: x ** e, g, Cc())
        op = p.ops[-1]
        assert op[0] == P[want[0]], e
        if want[0] == "POWI":
            assert op[2] == want[1]


def test_discrete_loads_and_nested_function_fields():
    """S = B + H (sia.jl:38-41) and ∂x S: loads at offsets, B inlined at the
    shifted centre coordinates, the reach of the loads recorded."""
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(8, 6))
    H = chmy.Field(None, g, Cc())
    B = chmy.FunctionField(lambda x, y: x * y, g, Cc())
    S = chmy.FunctionField(lambda g_, l, ix, iy, B, H: B[ix, iy] + H[ix, iy], g, Cc(), discrete=True,
                           parameters=(B, H))
    p = trace(lambda g_, l, ix, iy, S: S[ix, iy] - S[ix - 1, iy], g, (Vv(), Cc()), discrete=True, parameters=(S,))
    loads = [o for o in p.ops if o[0] == P["LOAD"]]
    assert len(p.srcs) == 1 and p.srcs[0] is H
    assert sorted(o[2] & 0xFF for o in loads) == [0, 0xFF]       # offsets 0 and -1 along x
    coords = [o for o in p.ops if o[0] == P["COORD"]]
    assert {(o[1], o[2], o[3]) for o in coords} == {(0, 0, 0), (0, 0, -1), (1, 0, 0)}   # B at centres
    assert p.reach == (1, 0)
    assert S.depends_on_fields() and not B.depends_on_fields()


def test_untraceable_functions_raise():
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(4, 4))
    with pytest.raises(TraceError):
        trace(lambda x, y: x if x > 0.5 else y, g, Cc())
    with pytest.raises(TraceError):
        trace(lambda x, y: min(x, y), g, Cc())
    with pytest.raises(TraceError):
        trace(lambda x, y: np.arctan2(x, y), g, Cc())
    H = chmy.Field(None, g, Cc())
    with pytest.raises(TraceError):
        trace(lambda g_, l, ix, iy, H: H[iy, ix], g, Cc(), discrete=True, parameters=(H,))


def test_host_field_set_evaluates_with_numpy():
    """On a host field (no GPU), set! evaluates the same function and the
    same FunctionField with numpy at the fma coordinates."""
    g = chmy.UniformGrid(origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(8, 6))
    f = chmy.Field(None, g, (Vv(), Cc()))
    fn = lambda x, y, a: a * x - y
    chmy.set_(f, fn, g, parameters=(3.0,))
    xs = np.meshgrid(g.coords(f.loc, 0), g.coords(f.loc, 1), indexing="ij")
    assert np.array_equal(f.interior().numpy(), 3.0 * xs[0] - xs[1])
    h = chmy.Field(None, g, (Vv(), Cc()))
    chmy.set_(h, chmy.FunctionField(fn, g, (Vv(), Cc()), parameters=dict(a=3.0)))
    assert np.array_equal(h.interior().numpy(), f.interior().numpy())


def _fdesc(n, loc, data=0x1000):
    f = _lib.ChmyField()
    f.data, f.nd, f.dtype = data, 2, _lib.F64
    for d in range(2):
        f.n[d], f.halo[d], f.loc[d] = n[d] + loc[d], 1, loc[d]
    return f


def test_library_rejects_malformed_programs_on_host():
    L = _lib.lib()
    g = chmy.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 1.0), dims=(8, 6)).desc()
    dst = _fdesc((8, 6), (0, 0))
    k = (C.c_double * 1)(1.0)

    def run(ops, srcs=None, nsrc=0, box=None):
        arr = (_lib.ChmyPop * len(ops))()
        for q, o in enumerate(ops):
            arr[q].op, arr[q].a, arr[q].b, arr[q].c = o
        return L.chmy_set_program(C.byref(g), C.byref(dst), box, arr, len(ops), k, 1, srcs, nsrc, None)

    # an operand that is not an earlier register
    assert run([(P["CONST"], 0, 0, 0), (P["ADD"], 0, 1, 0)]) == -1
    assert b"malformed" in L.chmy_last_error()
    # a constant slot past the table, a coordinate of a missing dim, a bad opcode
    assert run([(P["CONST"], 1, 0, 0)]) == -1
    assert run([(P["COORD"], 2, 0, 0)]) == -1
    assert run([(7, 0, 0, 0)]) == -1
    # too long
    assert run([(P["CONST"], 0, 0, 0)] * 129) == -7
    # a load that leaves the source's storage over the box
    src = (_lib.ChmyField * 1)(_fdesc((8, 6), (0, 0), 0x2000))
    assert run([(P["LOAD"], 0, 3, 0)], src, 1) == -4          # x offset +3 past n + 2H
    assert b"storage" in L.chmy_last_error()
    # a box outside dst's storage
    b = _lib.ChmyBox()
    b.lo[0], b.lo[1], b.len[0], b.len[1] = -2, 1, 4, 4
    assert run([(P["CONST"], 0, 0, 0)], box=C.byref(b)) == -4
