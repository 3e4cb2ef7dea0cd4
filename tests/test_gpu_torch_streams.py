"""The device entry points on a caller's stream: torch tensors as the device
arrays and a torch.cuda.Stream as the stream, with torch work ordered before
and after the library's launches on that same stream (no host
synchronisation in between).  Results against the oracle."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_execute_dct_and_batch_on_a_torch_stream(amd, orc):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N = 1 << 16
    announce(f"torch stream N={N}")
    u = amd.coefficients(N, seed=21)
    plan = amd.FMMPlan(N, 64, 18, amd.BOTH)
    dct = amd.DCTPlan(N)
    bp = amd.BatchPlan([100, 3000, 512], 18, amd.L2C, tune=False)
    ub = amd.packed_coefficients([100, 3000, 512], seed=22)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        x = torch.zeros(N, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        v = torch.empty_like(x)
        xb = torch.zeros(ub.size, dtype=torch.float64, device="cuda")
        yb = torch.empty_like(xb)
        # the inputs are written by torch on s, then used by the library on s
        x.copy_(torch.from_numpy(u), non_blocking=False)
        x.mul_(2.0)
        xb.copy_(torch.from_numpy(ub))
        sp = s.cuda_stream
        plan.execute_device(x.data_ptr(), y.data_ptr(), amd.L2C, 1, sp)
        dct.execute_device(y.data_ptr(), v.data_ptr(), 3, sp)
        bp.execute_device(xb.data_ptr(), yb.data_ptr(), amd.L2C, sp)
        yb.mul_(0.5)  # torch work after the batch, same stream
        out_y, out_v, out_b = y.cpu(), v.cpu(), yb.cpu()
    s.synchronize()
    ref = orc.FMMPlan(N, 64, 18, orc.BOTH).execute(2 * u, 0)
    tol = 1e-14 * math.log2(N)
    assert relerr(out_y.numpy(), ref) < tol
    assert relerr(out_v.numpy(), orc.redft(ref, 3)) < 1e-15 * math.log2(2 * N) + tol
    off = 0
    for n in (100, 3000, 512):
        r = 0.5 * orc.FMMPlan(n, 64, 18, orc.BOTH).execute(ub[off:off + n], 0)
        assert relerr(out_b.numpy()[off:off + n], r) < 1e-14 * math.log2(n)
        off += n
