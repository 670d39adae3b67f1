"""chmy_probe_stream (the access-pattern probe behind bench.py --pattern-probe)
writes, in every pattern, the sum of the values it reads plus the output's
index on the Launcher box [0, n+1]^3 and nothing outside it."""
import ctypes as C

import numpy as np
import pytest
import torch

from chmy import _lib
from chmy.architectures import current_stream

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pattern,by,kchunk,xcd", [(0, 12, 16, 1), (1, 12, 16, 1), (1, 8, 5, 0), (1, 16, 3, 1),
                                                   (2, 12, 1, 1), (2, 4, 1, 0), (3, 12, 1, 1), (3, 8, 1, 0)])
@pytest.mark.parametrize("n", [20, 61])
def test_probe_writes_the_sum_on_the_box(n, pattern, by, kchunk, xcd):
    nr, nw = 18, 10
    size = (n + 4) ** 3
    rng = np.random.default_rng(n + pattern)
    ins = [torch.from_numpy(rng.uniform(-1, 1, size)).cuda() for _ in range(nr)]
    outs = [torch.zeros(size, dtype=torch.float64, device="cuda") for _ in range(nw)]
    pin = (C.c_void_p * nr)(*[t.data_ptr() for t in ins])
    pout = (C.c_void_p * nw)(*[t.data_ptr() for t in outs])
    _lib.call("chmy_probe_stream", pattern, pin, nr, pout, nw, n, by, kchunk, xcd, current_stream())
    torch.cuda.synchronize()
    A = [t.cpu().numpy().reshape((n + 4,) * 3) for t in ins]
    s = sum(A)
    if pattern == 3:   # plus every input at plane z-1 and the first three at z+1
        s = s + sum(np.roll(x, 1, axis=0) for x in A) + sum(np.roll(x, -1, axis=0) for x in A[:3])
    box = (slice(1, n + 3),) * 3     # logical 0 .. n+1 at storage 1 .. n+2
    inside = np.zeros((n + 4,) * 3, dtype=bool)
    inside[box] = True
    for q, t in enumerate(outs):
        o = t.cpu().numpy().reshape((n + 4,) * 3)
        np.testing.assert_allclose(o[inside], (s + q)[inside], rtol=1e-12, atol=1e-12)
        assert not o[~inside].any()
