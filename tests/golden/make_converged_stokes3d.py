"""Writes stokes3d_inclusion_24_converged.npz: the oracle's converged state of
the 3D PT viscoelastic Stokes inclusion workload at 24^3 (the benchmark's
setup scaled down: [-0.5, 0.5]^3, eta = (|x| < 0.1 ? 1e-3 : 1) exp(-T),
G = 1, dt from the diffusion limit, pure-shear free-slip walls +-1).

The oracle runs the reference's sequence (update_stress!, update_velocity!,
the V batch; examples/stokes_3d_inc_ve_T.jl:164-180) from rest until the
residual max(|divV|, |r_V|) checked every 200 iterations is below 1e-8;
about 83,000 iterations, a few minutes on 8 cores.  Stored: the interiors of
V, tau, Pr (x fastest, as the storage), the iteration count and the final
residual.  Run from the repository root:  python3 tests/golden/make_converged_stokes3d.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "chmy.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import oracle as O                                    # noqa: E402
from chmy import configs                              # noqa: E402
from chmy.stokes import pure_shear_values             # noqa: E402
from test_oracle_stokes_adjoint import Problem       # noqa: E402

N, EPS = 24, 1e-8


def main():
    g, eta, dt = configs.stokes3d_inclusion(None, N)
    P = Problem((N,) * 3)
    P.eta.arr[...] = eta
    O.maxloc(P.g, P.eml, P.eta)
    P.dt = dt
    it = P.solve(pure_shear_values(3), eps=EPS)
    out = {"iterations": np.int64(it), "eps": np.float64(EPS), "n": np.int64(N), "dt": np.float64(dt)}
    inner = lambda f: np.ascontiguousarray(f.arr[(slice(2, -2),) * 3])
    for D in range(3):
        out[f"V{D}"] = inner(P.V[D])
    for c in range(6):
        out[f"tau{c}"] = inner(P.tau[c])
    out["Pr"] = inner(P.Pr)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stokes3d_inclusion_24_converged.npz")
    np.savez_compressed(path, **out)
    print(f"{path}: {it} iterations")


if __name__ == "__main__":
    main()
