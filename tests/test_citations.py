"""Reference citations in the committed sources point at lines that exist and,
where a citation names a reference kernel, at that kernel.

Every `name.jl:a-b[,c-d]` in the product sources, the header, the oracle,
the tests and the docs is resolved against the reference checkout (when it is
present; the GPU host has none, so this test skips there).  Checks:
- the file exists (matched by its trailing path) and each range lies inside it;
- a citation on a line that names a kernel defined in the cited file
  (`update_stress!`, `compute_q!`, ...) overlaps that kernel's definition or
  holds a call to it;
- the anchors below: ranges the C header and the solver cite for specific
  code.
"""
import os
import re
import subprocess

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")

CITE = re.compile(r"([A-Za-z0-9_./-]+\.jl):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")
KERNEL = re.compile(r"\b(update_[A-Za-z_]+!|compute_q!|bc_kernel!)")
SKIP = ("SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md", "BASELINE.json")

# (file, first line, last line) -> text that must occur in that range
ANCHORS = [
    ("examples/stokes_3d_inc_ve_T.jl", 11, 21, "function update_old!"),
    ("examples/stokes_3d_inc_ve_T.jl", 23, 46, "function update_stress!"),
    ("examples/stokes_3d_inc_ve_T.jl", 48, 57, "function update_velocity!"),
    ("examples/stokes_3d_inc_ve_T.jl", 59, 71, "function update_thermal_flux!"),
    ("examples/stokes_3d_inc_ve_T.jl", 73, 77, "function update_thermal!"),
    ("examples/stokes_3d_inc_ve_T.jl", 105, 112, "dτ_Pr   = r / θ_dτ"),
    ("examples/stokes_3d_inc_ve_T.jl", 130, 132, "V.z => (x=Neumann(), y=Neumann(), z=Dirichlet())"),
    ("examples/stokes_3d_inc_ve_T.jl", 164, 165, "update_velocity! => (V, r_V, Pr, τ, ρgz, η_ve, νdτ, grid); bc=batch"),
    ("examples/stokes_3d_inc_ve_T.jl", 171, 175, "bc!(arch, grid, r_V.x => (x=Dirichlet(),)"),
    ("examples/stokes_3d_inc_ve_T.jl", 172, 175, "maximum(abs.(interior(r_V.z)))"),
    ("examples/stokes_3d_inc_ve_T.jl", 157, 157, "dt_diff = min(dx, dy, dz)^2 / λ_ρCp"),
    ("examples/stokes_2d_inc_ve_T.jl", 39, 40, "r_V.y[I...] = -∂y(Pr, g, I...)"),
    ("examples/sia.jl", 46, 49, "(H[ix, iy]^(n + 2) + Aₛ * H[ix, iy]^n) * vmag(∇S, g, ix, iy)^(n - 1)"),
    ("examples/sia.jl", 73, 73, "bc!(arch, grid, q.x => (x = Dirichlet()), q.y => (y = Dirichlet()))"),
    ("examples/diffusion_2d.jl", 8, 13, "function compute_q!"),
    ("examples/diffusion_2d.jl", 15, 19, "function update_C!"),
    ("src/KernelLaunch.jl", 41, 41, "worksize = size(grid, Center()) .+ 2"),
    ("src/KernelLaunch.jl", 117, 117, "KernelAbstractions.synchronize(backend)"),
    ("src/BoundaryConditions/batch.jl", 20, 29, "function bc!(arch::Architecture, grid::SG{N}, batch::BatchSet{N})"),
    ("src/BoundaryConditions/boundary_function.jl", 78, 81, "function FirstOrderBC{Kind}(f::F)"),
    ("src/Fields/field.jl", 59, 59, "KernelAbstractions.zeros(backend, type, data_size)"),
]


def _ref_files():
    out = {}
    for dp, _dn, fn in os.walk(REF):
        for f in fn:
            if f.endswith(".jl"):
                rel = os.path.relpath(os.path.join(dp, f), REF)
                out.setdefault(f, []).append(rel)
    return out


def _lines(rel):
    with open(os.path.join(REF, rel), encoding="utf-8") as fh:
        return fh.read().split("\n")


def _tracked():
    r = subprocess.run(["git", "ls-files"], cwd=ROOT, capture_output=True, text=True, check=True)
    return [p for p in r.stdout.split() if not p.startswith(SKIP) and p.endswith(
        (".py", ".hip", ".hpp", ".h", ".cpp", ".md", ".jl", ".txt")) and os.path.isfile(os.path.join(ROOT, p))]


def _resolve(files, name):
    cands = [c for c in files.get(os.path.basename(name), []) if c.endswith(name) or "/" not in name]
    return cands


def _definition(L, k):
    """(first, last) line of `function k(` ... `end` in L, or None."""
    for i, line in enumerate(L):
        if f"function {k}(" in line:
            for j in range(i + 1, len(L)):
                if L[j] == "end":
                    return i + 1, j + 1
            return i + 1, len(L)
    return None


def _ranges(spec):
    for part in spec.split(","):
        a, _, b = part.strip().partition("-")
        yield int(a), int(b or a)


def test_citations_resolve_and_cover_named_kernels():
    files = _ref_files()
    bad = []
    for path in _tracked():
        with open(os.path.join(ROOT, path), encoding="utf-8", errors="replace") as fh:
            text = fh.read().split("\n")
        for ln, line in enumerate(text, 1):
            for m in CITE.finditer(line):
                name, spec = m.group(1), m.group(2)
                cands = _resolve(files, name)
                if not cands:
                    bad.append(f"{path}:{ln}: {name} does not exist in the reference")
                    continue
                L = _lines(cands[0])
                for a, b in _ranges(spec):
                    if a < 1 or b < a or b > len(L):
                        bad.append(f"{path}:{ln}: {name}:{a}-{b} outside its {len(L)} lines")
                        continue
                    body = "\n".join(L[a - 1:b])
                    for k in KERNEL.findall(line):
                        span = _definition(L, k)
                        if span and k not in body and not (a <= span[1] and b >= span[0]):
                            bad.append(f"{path}:{ln}: {name}:{a}-{b} neither defines nor calls {k}")
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("rel,a,b,text", ANCHORS)
def test_anchor(rel, a, b, text):
    L = _lines(rel)
    assert b <= len(L)
    assert text in "\n".join(L[a - 1:b]), f"{rel}:{a}-{b} does not hold {text!r}"
