"""The Julia binding (chmy.jl_amd/julia/ChmyMI355X.jl) against the C ABI.

The image has no Julia, so the binding cannot run here.  These CPU checks keep
it from drifting from include/chmy_mi355x.h: every `ccall` names a declared
entry point with the declared number of arguments and matching types, passes
as many arguments as it declares types, the descriptor structs have the C
layouts, the enum constants have the C values, every hot-path kernel of the
reference examples has a binding, and the file's blocks and brackets balance.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "chmy.jl_amd", "julia", "ChmyMI355X.jl")
HDR = os.path.join(ROOT, "include", "chmy_mi355x.h")

# C parameter type -> the Julia ccall types that pass it
C2JL = {
    "const chmy_grid_t*": {"Ref{ChmyGrid}", "Ptr{ChmyGrid}"},
    "const chmy_box_t*": {"Ref{ChmyBox}", "Ptr{ChmyBox}"},
    "const chmy_field_t*": {"Ref{ChmyField}", "Ptr{ChmyField}"},
    "const chmy_bc_t*": {"Ref{ChmyBC}", "Ptr{ChmyBC}"},
    "void*": {"Ptr{Cvoid}"},
    "const void*": {"Ptr{Cvoid}"},
    "void**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "int": {"Cint"},
    "double": {"Cdouble"},
    "size_t": {"Csize_t"},
    "int64_t": {"Int64"},
    "const int32_t*": {"Ptr{Int32}"},
    "int*": {"Ref{Cint}", "Ptr{Cint}"},
    "double*": {"Ref{Cdouble}", "Ptr{Cdouble}"},
    "const double*": {"Ptr{Cdouble}"},
}
RET2JL = {"int": "Cint", "const char*": "Cstring"}


def _strip_c_comments(s):
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def header_prototypes():
    src = _strip_c_comments(open(HDR).read())
    protos = {}
    for m in re.finditer(r"(int|const char\s*\*)\s*(chmy_\w+)\s*\(([^;{]*?)\)\s*;", src):
        ret = re.sub(r"\s+", "", m.group(1)).replace("char*", "char*").replace("constchar*", "const char*")
        params = []
        body = m.group(3).strip()
        if body and body != "void":
            for p in body.split(","):
                p = " ".join(p.split())
                # drop the parameter name: the last identifier
                t = re.sub(r"\s*\b\w+\s*$", "", p).strip()
                t = re.sub(r"\s*\*", "*", t)
                params.append(t)
        protos[m.group(2)] = (ret, params)
    return protos


def header_structs():
    src = _strip_c_comments(open(HDR).read())
    out = {}
    for m in re.finditer(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            mm = re.match(r"(.*?)\s*(\w+)\s*(\[(\d+)\])?$", decl)
            typ = re.sub(r"\s*\*", "*", mm.group(1).strip())
            for name in [mm.group(2)]:
                fields.append((name, typ, int(mm.group(4)) if mm.group(4) else None))
        out[m.group(2)] = fields
    return out


def header_enums():
    src = _strip_c_comments(open(HDR).read())
    vals = {}
    for m in re.finditer(r"enum\s*\{(.*?)\}\s*;", src, flags=re.S):
        for item in m.group(1).split(","):
            mm = re.match(r"\s*(CHMY_\w+)\s*=\s*(-?\d+)", item)
            if mm:
                vals[mm.group(1)] = int(mm.group(2))
    return vals


def julia_source():
    return open(JL).read()


def _strip_julia(src):
    """Source with strings, character literals and comments blanked out."""
    out = []
    i, n = 0, len(src)
    while i < n:
        if src.startswith('"""', i):
            j = src.index('"""', i + 3) + 3
            out.append(" " * (j - i))
            i = j
        elif src[i] == '"':
            j = i + 1
            while src[j] != '"':
                j += 2 if src[j] == "\\" else 1
            out.append(" " * (j + 1 - i))
            i = j + 1
        elif src.startswith("#=", i):
            j = src.index("=#", i) + 2
            out.append(" " * (j - i))
            i = j
        elif src[i] == "#":
            j = src.find("\n", i)
            j = n if j < 0 else j
            out.append(" " * (j - i))
            i = j
        else:
            out.append(src[i])
            i += 1
    return "".join(out)


def _split_top(s):
    """Split s (the inside of a bracket pair) at top-level commas."""
    parts, depth, cur = [], 0, []
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        parts.append("".join(cur).strip())
    return parts


def _matching(s, i):
    """Index of the bracket closing the one at s[i]."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] in "([{":
            depth += 1
        elif s[j] in ")]}":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced")


def julia_ccalls():
    src = _strip_julia(julia_source())
    calls = []
    for m in re.finditer(r"\bccall\(", src):
        o = m.end() - 1
        c = _matching(src, o)
        args = _split_top(src[o + 1:c])
        target = args[0]
        tm = re.match(r"\(\s*:(\w+)\s*,\s*LIB\s*\)$", target)
        assert tm, f"ccall target {target!r} is not (:name, LIB)"
        ret = args[1]
        types = args[2].strip()
        assert types.startswith("(") and types.endswith(")"), types
        tlist = _split_top(types[1:-1])
        calls.append((tm.group(1), ret, tlist, args[3:], src.count("\n", 0, m.start()) + 1))
    return calls


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 25
    for name, ret, types, args, line in calls:
        assert name in protos, f"line {line}: {name} is not declared in chmy_mi355x.h"
        cret, cparams = protos[name]
        assert ret == RET2JL[cret], f"line {line}: {name} returns {cret}, binding says {ret}"
        assert len(types) == len(cparams), f"line {line}: {name} takes {len(cparams)} arguments, binding declares {len(types)}"
        assert len(args) == len(types), f"line {line}: {name} declares {len(types)} types but passes {len(args)}"
        for k, (jt, ct) in enumerate(zip(types, cparams)):
            assert ct in C2JL, f"{name}: no Julia mapping for C type {ct!r}"
            assert jt in C2JL[ct], f"line {line}: {name} argument {k + 1} is {ct}, binding passes {jt}"


def test_binding_covers_the_hot_path_entry_points():
    bound = {c[0] for c in julia_ccalls()}
    need = {"chmy_init", "chmy_malloc_zeros", "chmy_free", "chmy_memset_zero", "chmy_memcpy_h2d", "chmy_memcpy_d2h",
            "chmy_stream_create", "chmy_stream_sync", "chmy_last_error", "chmy_bc_batchset", "chmy_bc_side",
            "chmy_diffusion_flux", "chmy_diffusion_update", "chmy_stokes_update_old", "chmy_stokes_stress",
            "chmy_stokes_velocity", "chmy_thermal_flux", "chmy_thermal_update", "chmy_sia_flux", "chmy_sia_update",
            "chmy_stokes_iterate", "chmy_sia_run", "chmy_sia_reverse", "chmy_stokes_grad_eta",
            "chmy_stokes_iter2", "chmy_stokes_iter2_first", "chmy_stokes_residual", "chmy_stokes_accumulate",
            "chmy_maxabs", "chmy_op", "chmy_op_masked"}
    assert need <= bound, f"entry points without a binding: {sorted(need - bound)}"


def test_every_example_kernel_has_a_binding():
    src = _strip_julia(julia_source())
    kernels = ["compute_q!", "update_C!", "update_old!", "update_stress!", "update_velocity!",
               "update_thermal_flux!", "update_thermal!", "update_q!", "update_H!"]
    for k in kernels:
        assert f"const {k} = HipKernel{{:{k}}}()" in src, k
        assert re.search(r"function run_kernel\(::HipKernel\{:" + re.escape(k) + r"\}", src), k


C2JL_FIELD = {"void*": "Ptr{Cvoid}", "int32_t": "Int32", "int64_t": "Int64", "double": "Float64",
              "const chmy_field_t*": "Ptr{ChmyField}", "const double*": "Ptr{Float64}"}


@pytest.mark.parametrize("cname,jname", [("chmy_field_t", "ChmyField"), ("chmy_grid_t", "ChmyGrid"),
                                         ("chmy_box_t", "ChmyBox"), ("chmy_bc_t", "ChmyBC")])
def test_descriptor_layouts(cname, jname):
    cfields = header_structs()[cname]
    src = _strip_julia(julia_source())
    m = re.search(r"\bstruct " + jname + r"\n(.*?)\nend", src, flags=re.S)
    assert m, jname
    jfields = [tuple(x.strip() for x in line.split("::")) for line in m.group(1).strip().splitlines()]
    assert len(jfields) == len(cfields), (jfields, cfields)
    for (jn, jt), (cn, ct, cnt) in zip(jfields, cfields):
        assert jn == cn, (jname, jn, cn)
        base = C2JL_FIELD[ct]
        want = f"NTuple{{{cnt},{base}}}" if cnt else base
        assert jt == want, (jname, jn, jt, want)


def test_constants_match_the_enums():
    enums = header_enums()
    src = _strip_julia(julia_source())
    found = 0
    for m in re.finditer(r"const (CHMY_\w+) = (?:Int32|Cint)\((-?\d+)\)", src):
        assert enums[m.group(1)] == int(m.group(2)), m.group(1)
        found += 1
    assert found >= 8


def test_blocks_and_brackets_balance():
    src = _strip_julia(julia_source())
    depth = 0
    for ch in src:
        depth += ch in "([{"
        depth -= ch in ")]}"
        assert depth >= 0
    assert depth == 0
    # block keywords outside brackets against `end` outside brackets
    openers = re.compile(r"\b(module|function|macro|struct|if|for|while|let|begin|do|try|quote)\b")
    level, stack, i = 0, 0, 0
    depth = 0
    toks = re.finditer(r"[()\[\]{}]|\b(?:module|function|macro|struct|if|for|while|let|begin|do|try|quote|end|elseif|else)\b", src)
    for t in toks:
        s = t.group(0)
        if s in "([{":
            depth += 1
        elif s in ")]}":
            depth -= 1
        elif depth == 0 and openers.fullmatch(s):
            # `mutable struct` counts once (struct); one-line `f(x) = ...` has no keyword
            level += 1
        elif depth == 0 and s == "end":
            level -= 1
            assert level >= 0, src.count("\n", 0, t.start()) + 1
    assert level == 0


def test_no_reference_method_is_replaced():
    """Every method the binding adds to a Chmy or KernelAbstractions function
    dispatches on one of its own types (MI355XBackend, MI355XDevice,
    MI355XArray, MI355XField, MI355XArch), so it replaces no existing method."""
    src = _strip_julia(julia_source())
    own = ("MI355XBackend", "MI355XDevice", "MI355XArray", "MI355XField", "MI355XArch")
    for m in re.finditer(r"^(?:function )?((?:KernelAbstractions\.|Base\.)?(?:get_device|set_device!|heuristic_groupsize|"
                         r"pointertype|gpu_aware_compat|bc!|set!|zeros|allocate|synchronize|get_backend|priority!|"
                         r"unsafe_wrap|fill!|copyto!|Array))\((.*)$", src, flags=re.M):
        sig = m.group(2)
        assert any(o in sig for o in own), f"{m.group(1)}({sig[:60]}...) does not dispatch on a binding type"
