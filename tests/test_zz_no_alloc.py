"""chmy_sia_reverse_noalloc, the reverse sweep that keeps its segment states
and second adjoint buffer in the library's file-scope device scratch, against
chmy_sia_reverse with caller-allocated, zero-filled workspace: dJ/dH_0 and
dJ/dlog A bitwise equal, on two calls in a row, the second on a smaller grid
with a short last segment (so the scratch holds the first call's data when
the second starts)."""
import ctypes as C

import numpy as np
import pytest
import torch

import chmy
from chmy import _lib, configs
from chmy.architectures import current_stream
from chmy.kernels import _arr
from chmy.sia import SIA

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arch():
    return chmy.Arch(chmy.MI355XBackend())


def _reverse_both(arch, n, nsteps, ncheck):
    """Checkpointed forward run, then both reverse sweeps from the same
    checkpoints and dJ/dH_N; returns (lam, grad) of each."""
    g, B, la0, la1, mb, dt = configs.sia_ice_cap(arch, n)
    dts = [dt] * nsteps
    S = SIA(arch, g, B, logA=la1, As=0.0, npow=3, mb=mb)
    S.reset()
    S.forward(nsteps, dts)
    H_obs = S.H.interior().clone()
    S.set_logA(la0)
    # chmy_sia_reverse, caller-allocated seg and lam_b (zero-filled)
    J, grad = S.gradient(None, H_obs, nsteps, dts, ncheck)
    lam_ref = S.lam.data.clone()
    grad_ref = S.grad.data.clone()
    # the same checkpoints, the no-allocation entry point
    lam = S._field()
    lam.interior().copy_(S.H.interior() - H_obs)
    gA = S._field()
    d = (C.c_double * nsteps)(*dts)
    print(f"chmy_sia_reverse_noalloc n={n} nsteps={nsteps} stride={S.stride}", flush=True)
    _lib.call("chmy_sia_reverse_noalloc", C.byref(g.desc()), _arr(S.ckpt), len(S.ckpt), S.stride, nsteps,
              C.byref(lam.desc()), C.byref(gA.desc()), *S._args(), d, current_stream())
    torch.cuda.synchronize()
    return (lam_ref, grad_ref), (lam.data, gA.data)


def test_noalloc_reverse_equals_workspace_reverse(arch):
    # the configured workload: 512^2 cells, 300 steps, 30 checkpoints (stride 10)
    (l0, g0), (l1, g1) = _reverse_both(arch, 512, 300, 30)
    assert torch.equal(l1, l0)
    assert torch.equal(g1, g0)
    assert float(g0.abs().max()) > 0
    # then a smaller grid: 96^2, 23 steps at stride 5 (last segment 3 steps)
    (l0, g0), (l1, g1) = _reverse_both(arch, 96, 23, 5)
    assert torch.equal(l1, l0)
    assert torch.equal(g1, g0)
    assert float(g0.abs().max()) > 0


def test_noalloc_reverse_refuses_oversize(arch):
    """1024^2 cells at stride 10 need 19 Center storages of 8.45 MB: above
    the 48 MiB scratch, refused with CHMY_E_SIZE before anything runs."""
    g, B, la0, la1, mb, dt = configs.sia_ice_cap(arch, 1024)
    S = SIA(arch, g, B, logA=la0, As=0.0, npow=3, mb=mb)
    ck = [S._field() for _ in range(3)]
    d = (C.c_double * 30)(*([dt] * 30))
    with pytest.raises(_lib.ChmyError, match="scratch"):
        _lib.call("chmy_sia_reverse_noalloc", C.byref(g.desc()), _arr(ck), 3, 10, 30, C.byref(S.lam.desc()),
                  C.byref(S.grad.desc()), *S._args(), d, current_stream())
