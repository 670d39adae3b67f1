"""FunctionField and set! on the device (src/Fields/function_field.jl,
src/Fields/field.jl:119-142).

The reference hands the user's function to KernelAbstractions, which compiles
it into a GPU kernel.  Here the function is traced once on the host: it is
called with symbolic stand-ins for the coordinates (continuous) or the
logical indices (discrete), and every arithmetic operation, numpy ufunc,
comparison and `ifelse` it applies is recorded as one operation of a point
program.  `chmy_set_program` then runs that program at every point on the
GPU.  The same Python function still runs on numpy arrays, so one definition
serves the host and the device.

What a traced function may do: + - * / ** (an int exponent -2..3 is Julia's
literal power, repeated products), unary -, abs, comparisons, & | ~ on
comparison results, `ifelse(c, a, b)`, numpy's exp, log, sqrt, sin, cos,
tanh, floor, minimum, maximum, power, square, and, in discrete functions,
index a parameter Field or FunctionField at the point's indices plus integer
offsets (`H[ix, iy]`, `S[ix + 1, iy]`).  A Python `if` on a traced value, or
`min`/`max` builtins, raise TraceError: use `ifelse` and `np.minimum`.
"""
from __future__ import annotations

import ctypes as C
import math
import numbers

import numpy as np
import torch

from . import _lib
from .architectures import current_stream
from .grids import expand_loc

MAX_OPS, MAX_CONST, MAX_SRC = 128, 64, 8

P = {name: getattr(_lib, "P_" + name) for name in (
    "CONST", "INDEX", "COORD", "SPACING", "INV_SPACING", "LOAD", "NEG", "ABS", "SQRT", "EXP", "LOG", "SIN", "COS",
    "TANH", "FLOOR", "NOT", "ADD", "SUB", "MUL", "DIV", "POW", "MIN", "MAX", "LT", "LE", "GT", "GE", "EQ", "NE",
    "AND", "OR", "POWI", "SELECT", "FMA")}


class TraceError(TypeError):
    pass


class _Tracer:
    """Builds one point program: operations deduplicated (a repeated
    subexpression is computed once), constants and source fields in tables."""

    def __init__(self, grid):
        self.grid = grid
        self.ops = []
        self.memo = {}
        self.consts = []
        self.cmemo = {}
        self.srcs = []
        self.reach = [0] * grid.nd      # largest |load offset| per dim

    def emit(self, op, a=0, b=0, c=0):
        key = (op, a, b, c)
        r = self.memo.get(key)
        if r is None:
            if len(self.ops) >= MAX_OPS:
                raise TraceError(f"point program longer than {MAX_OPS} operations")
            r = len(self.ops)
            self.ops.append(key)
            self.memo[key] = r
        return Expr(self, r)

    def const(self, v):
        v = float(v)
        k = (v, math.copysign(1.0, v)) if v == v else ("nan",)
        slot = self.cmemo.get(k)
        if slot is None:
            if len(self.consts) >= MAX_CONST:
                raise TraceError(f"more than {MAX_CONST} constants")
            slot = len(self.consts)
            self.consts.append(v)
            self.cmemo[k] = slot
        return self.emit(P["CONST"], slot)

    def src(self, field):
        for q, f in enumerate(self.srcs):
            if f is field:
                return q
        if len(self.srcs) >= MAX_SRC:
            raise TraceError(f"more than {MAX_SRC} source fields")
        self.srcs.append(field)
        return len(self.srcs) - 1

    def expr(self, v):
        if isinstance(v, Expr):
            if v.t is not self:
                raise TraceError("a traced value from another trace")
            return v
        if isinstance(v, Index):
            return v.expr()
        if isinstance(v, (bool, np.bool_)):
            return self.const(1.0 if v else 0.0)
        if isinstance(v, numbers.Real):
            return self.const(v)
        raise TraceError(f"cannot use {type(v).__name__} in a traced function")


_UNARY = {np.negative: "NEG", np.absolute: "ABS", np.fabs: "ABS", np.sqrt: "SQRT", np.exp: "EXP", np.log: "LOG",
          np.sin: "SIN", np.cos: "COS", np.tanh: "TANH", np.floor: "FLOOR", np.logical_not: "NOT"}
_BINARY = {np.add: "ADD", np.subtract: "SUB", np.multiply: "MUL", np.true_divide: "DIV", np.minimum: "MIN",
           np.maximum: "MAX", np.less: "LT", np.less_equal: "LE", np.greater: "GT", np.greater_equal: "GE",
           np.equal: "EQ", np.not_equal: "NE", np.logical_and: "AND", np.logical_or: "OR"}


class Expr:
    """A traced f64 value: register `r` of tracer `t`."""
    __slots__ = ("t", "r")
    __array_priority__ = 1000
    __hash__ = None

    def __init__(self, t, r):
        self.t, self.r = t, r

    def _bin(self, name, a, b):
        t = self.t
        return t.emit(P[name], t.expr(a).r, t.expr(b).r)

    def __add__(self, o): return self._bin("ADD", self, o)
    def __radd__(self, o): return self._bin("ADD", o, self)
    def __sub__(self, o): return self._bin("SUB", self, o)
    def __rsub__(self, o): return self._bin("SUB", o, self)
    def __mul__(self, o): return self._bin("MUL", self, o)
    def __rmul__(self, o): return self._bin("MUL", o, self)
    def __truediv__(self, o): return self._bin("DIV", self, o)
    def __rtruediv__(self, o): return self._bin("DIV", o, self)
    def __lt__(self, o): return self._bin("LT", self, o)
    def __le__(self, o): return self._bin("LE", self, o)
    def __gt__(self, o): return self._bin("GT", self, o)
    def __ge__(self, o): return self._bin("GE", self, o)
    def __eq__(self, o): return self._bin("EQ", self, o)
    def __ne__(self, o): return self._bin("NE", self, o)
    def __and__(self, o): return self._bin("AND", self, o)
    def __rand__(self, o): return self._bin("AND", o, self)
    def __or__(self, o): return self._bin("OR", self, o)
    def __ror__(self, o): return self._bin("OR", o, self)
    def __invert__(self): return self.t.emit(P["NOT"], self.r)
    def __neg__(self): return self.t.emit(P["NEG"], self.r)
    def __pos__(self): return self
    def __abs__(self): return self.t.emit(P["ABS"], self.r)

    def __pow__(self, o):
        if isinstance(o, (int, np.integer)) and not isinstance(o, bool) and -2 <= int(o) <= 3:
            return self.t.emit(P["POWI"], self.r, int(o))
        return self._bin("POW", self, o)

    def __rpow__(self, o):
        return self._bin("POW", o, self)

    def __bool__(self):
        raise TraceError("a traced value decides a Python branch: use chmy.ifelse (and np.minimum/np.maximum "
                         "instead of min/max)")

    def __float__(self):
        raise TraceError("a traced value has no host value")

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method != "__call__" or kwargs:
            return NotImplemented
        if ufunc in _UNARY:
            return self.t.emit(P[_UNARY[ufunc]], self.t.expr(inputs[0]).r)
        if ufunc is np.square:
            return self.t.emit(P["POWI"], self.t.expr(inputs[0]).r, 2)
        if ufunc is np.power:
            a, b = inputs
            if isinstance(a, Expr) or isinstance(a, Index):
                return self.t.expr(a) ** b
            return self._bin("POW", a, b)
        if ufunc in _BINARY:
            return self._bin(_BINARY[ufunc], *inputs)
        raise TraceError(f"numpy.{ufunc.__name__} is not available in a traced function")


class Index:
    """A traced logical index along dim d, plus a constant offset: indexes
    fields directly, and becomes an f64 value in arithmetic."""
    __array_priority__ = 1000

    def __init__(self, t, d, off=0):
        self.t, self.d, self.off = t, d, int(off)

    def expr(self):
        return self.t.emit(P["INDEX"], self.d, self.off)

    def __add__(self, o):
        if isinstance(o, (int, np.integer)) and not isinstance(o, bool):
            return Index(self.t, self.d, self.off + int(o))
        return self.expr() + o

    def __radd__(self, o):
        return self.__add__(o)

    def __sub__(self, o):
        if isinstance(o, (int, np.integer)) and not isinstance(o, bool):
            return Index(self.t, self.d, self.off - int(o))
        return self.expr() - o

    def __rsub__(self, o): return o - self.expr()
    def __mul__(self, o): return self.expr() * o
    def __rmul__(self, o): return o * self.expr()
    def __truediv__(self, o): return self.expr() / o
    def __rtruediv__(self, o): return o / self.expr()
    def __neg__(self): return -self.expr()
    def __pow__(self, o): return self.expr() ** o
    def __lt__(self, o): return self.expr() < o
    def __le__(self, o): return self.expr() <= o
    def __gt__(self, o): return self.expr() > o
    def __ge__(self, o): return self.expr() >= o

    def __bool__(self):
        raise TraceError("a traced index decides a Python branch: use chmy.ifelse")

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        inputs = tuple(x.expr() if isinstance(x, Index) else x for x in inputs)
        e = next(x for x in inputs if isinstance(x, Expr))
        return e.__array_ufunc__(ufunc, method, *inputs, **kwargs)


def _offsets(t, I, nd):
    if not isinstance(I, tuple):
        I = (I,)
    if len(I) != nd or not all(isinstance(i, Index) for i in I):
        raise TraceError("index fields with the traced indices plus integer offsets, one per dim")
    offs = []
    for d, i in enumerate(I):
        if i.d != d:
            raise TraceError(f"index {d} of a field is the traced index of dim {i.d}")
        if not -127 <= i.off <= 127:
            raise TraceError("index offset out of range")
        offs.append(i.off)
    return offs


class _FieldRef:
    """A Field parameter inside a discrete trace: f[ix + a, iy + b] loads it."""

    def __init__(self, t, field):
        self.t, self.field = t, field

    def __getitem__(self, I):
        offs = _offsets(self.t, I, self.t.grid.nd)
        packed = 0
        for d, o in enumerate(offs):
            packed |= (o & 0xFF) << (8 * d)
            self.t.reach[d] = max(self.t.reach[d], abs(o))
        return self.t.emit(P["LOAD"], self.t.src(self.field), packed)


class _FFRef:
    """A FunctionField parameter inside a trace: indexing it inlines its
    function at the shifted indices (the reference evaluates it lazily at
    each access, function_field.jl:59-61)."""

    def __init__(self, t, ff):
        self.t, self.ff = t, ff

    def __getitem__(self, I):
        offs = _offsets(self.t, I, self.t.grid.nd)
        return self.ff._trace_into(self.t, offs)


def ifelse(c, a, b):
    """Julia's ifelse: both branches evaluated, c selects (numpy arrays on the
    host, a SELECT operation in a traced function)."""
    if any(isinstance(x, (Expr, Index)) for x in (c, a, b)):
        t = next(x.t for x in (c, a, b) if isinstance(x, (Expr, Index)))
        return t.emit(P["SELECT"], t.expr(c).r, t.expr(a).r, t.expr(b).r)
    return np.where(c, a, b)


def _params(p):
    if p is None:
        return ()
    if isinstance(p, dict):
        return tuple(p.values())
    return tuple(p)


class Program:
    """A traced point program: operations, constants, source fields, and the
    largest load offset per dim."""

    def __init__(self, t, out):
        if isinstance(out, Index):
            out = out.expr()
        if not isinstance(out, Expr):
            out = t.expr(out)
        if out.r != len(t.ops) - 1:
            one = t.const(1.0)
            t.emit(P["SELECT"], one.r, out.r, out.r)   # the stored value is the last register
        self.ops = list(t.ops)
        self.consts = list(t.consts)
        self.srcs = list(t.srcs)
        self.reach = tuple(t.reach)
        self._c = None

    def __len__(self):
        return len(self.ops)

    def _cargs(self):
        if self._c is None:
            n = len(self.ops)
            ops = (_lib.ChmyPop * n)()
            for q, (op, a, b, c) in enumerate(self.ops):
                ops[q].op, ops[q].a, ops[q].b, ops[q].c = op, a, b, c
            k = (C.c_double * max(1, len(self.consts)))(*self.consts)
            self._c = (ops, n, k, len(self.consts))
        return self._c

    def run(self, grid, dst, lo=None, length=None):
        """Evaluate at every point of the box (logical indices of dst; default
        its interior) into dst."""
        ops, n, k, nk = self._cargs()
        box = None
        if lo is not None:
            box = _lib.ChmyBox()
            for d in range(grid.nd):
                box.lo[d], box.len[d] = lo[d], length[d]
        srcs = (_lib.ChmyField * max(1, len(self.srcs)))(*[f.desc() for f in self.srcs])
        _lib.call("chmy_set_program", C.byref(grid.desc()), C.byref(dst.desc()),
                  C.byref(box) if box is not None else None, ops, n, k, nk, srcs if self.srcs else None,
                  len(self.srcs), current_stream())


def trace(func, grid, loc, *, discrete=False, parameters=None):
    """The point program of `func` at location `loc`: continuous functions
    take (x, y, z, parameters...) at coord(grid, loc, I...), discrete ones
    (grid, loc, ix, iy, iz, parameters...) (function_field.jl:51-57)."""
    t = _Tracer(grid)
    locs = expand_loc(grid.nd, loc)
    out = _call(t, func, grid, locs, discrete, _params(parameters), [0] * grid.nd)
    return Program(t, out)


def _wrap_param(t, p):
    from .fields import Field
    if isinstance(p, Field):
        return _FieldRef(t, p)
    if isinstance(p, FunctionField):
        return _FFRef(t, p)
    return p


def _call(t, func, grid, locs, discrete, params, offs):
    wrapped = tuple(_wrap_param(t, p) for p in params)
    if discrete:
        idx = [Index(t, d, offs[d]) for d in range(grid.nd)]
        return func(grid, locs, *idx, *wrapped)
    xs = [t.emit(P["COORD"], d, locs[d].code, offs[d]) for d in range(grid.nd)]
    return func(*xs, *wrapped)


class FunctionField:
    """FunctionField(func, grid, loc; discrete=false, parameters=nothing)
    (function_field.jl:7-46): values computed from the function at each
    access.  On the device it is evaluated where it is read: inlined into the
    program of a set! or of another FunctionField that indexes it, or written
    by `materialize` into a Field for the Launcher kernels that take fields
    (ρg of update_velocity!, the SIA bed), before each launch that reads it
    when it depends on other fields, once when it depends only on the
    coordinates."""

    def __init__(self, func, grid, loc, *, discrete=False, parameters=None):
        self.func, self.grid = func, grid
        self.loc = expand_loc(grid.nd, loc)
        self.discrete = bool(discrete)
        self.parameters = parameters
        self.nd = grid.nd
        self.dims = grid.size(self.loc)
        self._prog = None
        self._cache = {}

    def size(self):
        return self.dims

    def location(self):
        return self.loc

    def program(self):
        if self._prog is None:
            self._prog = trace(self.func, self.grid, self.loc, discrete=self.discrete, parameters=self.parameters)
        return self._prog

    def depends_on_fields(self):
        return bool(self.program().srcs)

    def _trace_into(self, t, offs):
        return _call(t, self.func, self.grid, self.loc, self.discrete, _params(self.parameters), offs)

    def materialize(self, arch=None, dtype=torch.float64, out=None, halo=1):
        """A Field at this location holding the function's values over its
        whole storage (logical 1-2H .. n+2H), less the reach of any loads of
        other fields.  Cached when the values depend on coordinates only."""
        from .fields import Field
        device = arch.device if arch is not None else torch.device("cuda", torch.cuda.current_device())
        key = (dtype, halo, str(device))
        if out is None and not self.depends_on_fields() and key in self._cache:
            return self._cache[key]
        f = out if out is not None else Field(arch, self.grid, self.loc, dtype, halo=halo, device=device)
        prog = self.program()
        lo = tuple(1 - 2 * f.halo[d] + prog.reach[d] for d in range(self.nd))
        ln = tuple(f.dims[d] + 4 * f.halo[d] - 2 * prog.reach[d] for d in range(self.nd))
        prog.run(self.grid, f, lo, ln)
        if out is None and not self.depends_on_fields():
            self._cache[key] = f
        return f

    def interior(self, arch=None):
        """The values at the interior indices on the device (copy to the host
        with .cpu(), as Array(interior(f)) does)."""
        return self.materialize(arch).interior()

    def __getitem__(self, I):
        raise TraceError("a FunctionField is indexed inside a traced function (a parameter of another "
                         "FunctionField or set!), or read through materialize()/interior()")


def as_field(x, arch=None, dtype=torch.float64):
    """Fields pass through; a FunctionField is evaluated into one."""
    if isinstance(x, FunctionField):
        return x.materialize(arch, dtype)
    return x
