"""Converged PT Stokes solves of the benchmarked path (fast arithmetic form,
two iterations per pass, FP64) against the oracle's converged state.

The benchmark times single PT iterations, but what a user keeps is the
converged solution (examples/stokes_3d_inc_ve_T.jl:163-181: iterate until
max(|divV|, |r_V|) < eps).  The fast form differs from the reference's
expressions by a few ulp per iteration; these tests bound what that does to
the solution after the whole solve.

- 24^3 inclusion workload to 1e-8 from rest, against the oracle's converged
  state in tests/golden/stokes3d_inclusion_24_converged.npz (made by
  make_converged_stokes3d.py there: about 83,000 oracle iterations, minutes
  of CPU, too long to rerun in every test session).
- 63^3 inclusion workload to 1e-8: the fast form against the exact form,
  whose iterations are bitwise the oracle's sequence (test_gpu_fused.py,
  test_gpu_full_size.py).

Bar, per field: max|fast - reference| <= 1e-12 max|reference| (the FP64
parity bar; measured: at most 1.0e-14 at 24^3 and 1.9e-14 at 63^3, in Vy,
whose max is smallest), and the same iteration count (the checks fall every
200 iterations)."""
import os

import numpy as np
import pytest
import torch

import chmy
from chmy import configs
from chmy.stokes import PTStokes, pure_shear_values

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stokes3d_inclusion_24_converged.npz")
TOL = 1e-12


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _solve(arch, n, exact, eps=1e-8):
    g, eta, dt = configs.stokes3d_inclusion(arch, n)
    S = PTStokes(arch, g, eta_field=eta, G=1.0, dt=dt, normal_values=pure_shear_values(3), exact=exact)
    S.apply_bc()
    S.sync_buffers()
    it, err = S.solve(eps=eps, max_iter=400000, ncheck=200)
    torch.cuda.synchronize()
    assert max(err) < eps, err
    fields = {f"V{D}": S.V[D] for D in range(3)}
    fields.update({f"tau{c}": S.tau[c] for c in range(6)})
    fields["Pr"] = S.Pr
    return it, {k: f.xyz()[(slice(2, -2),) * 3].cpu().numpy() for k, f in fields.items()}


def _compare(got, ref):
    worst = {}
    for k, r in ref.items():
        scale = np.max(np.abs(r))
        assert scale > 0, k
        worst[k] = float(np.max(np.abs(got[k] - r)) / scale)
    print("max |difference| / max |field|:", {k: f"{v:.1e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= TOL, (k, v)


def test_fast_form_converged_solution_matches_the_oracle_24(arch):
    d = np.load(GOLDEN)
    n = int(d["n"])
    it, got = _solve(arch, n, exact=False, eps=float(d["eps"]))
    assert it == int(d["iterations"]), (it, int(d["iterations"]))
    ref = {k: np.transpose(d[k], (2, 1, 0)) for k in got}        # stored z, y, x; compared x, y, z
    _compare(got, ref)


def test_fast_form_converged_solution_matches_exact_form_63(arch):
    it_e, ref = _solve(arch, 63, exact=True)
    it_f, got = _solve(arch, 63, exact=False)
    assert it_f == it_e, (it_f, it_e)
    _compare(got, ref)
