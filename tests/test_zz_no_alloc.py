"""gm_inner_prod_noalloc (the inner product with its partial sums in the
library's static device scratch) against gm_inner_prod through the C ABI:
the same result bit for bit on two calls in a row, the second with a smaller
input, plus the reference's VectorTest known answer and an empty vector."""
import ctypes as C

import numpy as np
import pytest
import torch

from gpumatting_amd._lib import call, lib
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
S0 = C.c_void_p(0)


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_inner_prod_noalloc_matches_inner_prod(gpu):
    rng = np.random.default_rng(7)
    out = torch.zeros(4, dtype=torch.float32, device="cuda")
    for n in (3_000_001, 70_000, 1):  # second and third calls smaller
        x = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
        y = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
        call("gm_inner_prod", C.c_void_p(out.data_ptr()), _p(x), _p(y), n, S0)
        call("gm_inner_prod_noalloc", C.c_void_p(out.data_ptr() + 4), _p(x), _p(y), n, S0)
        torch.cuda.synchronize()
        r = out.cpu().numpy()
        assert r[0].view(np.uint32) == r[1].view(np.uint32), (n, r[0], r[1])
        want = O.dot_f32(x.cpu().numpy(), y.cpu().numpy())
        assert abs(float(r[1]) - want) <= 1e-6 * max(1.0, abs(want)), (n, r[1], want)
    # VectorTest.cu:36-65: <1, 2> over 1e6 elements
    x = torch.ones(1_000_000, dtype=torch.float32, device="cuda")
    call("gm_inner_prod_noalloc", C.c_void_p(out.data_ptr() + 8), _p(x), _p(2 * x), 1_000_000, S0)
    call("gm_inner_prod_noalloc", C.c_void_p(out.data_ptr() + 12), _p(x), _p(x), 0, S0)
    torch.cuda.synchronize()
    r = out.cpu().numpy()
    assert r[2] == 2e6 and r[3] == 0.0
    # argument errors come back as statuses, before any device work
    assert lib().gm_inner_prod_noalloc(None, _p(x), _p(x), 10, S0) == 1
    assert lib().gm_inner_prod_noalloc(C.c_void_p(out.data_ptr()), _p(x), _p(x), -1, S0) == 1
