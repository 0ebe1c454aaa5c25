"""ccd_mbpt2_noalloc (the entry point that allocates nothing) against
ccd_mbpt2 on the same contexts, through the C ABI: two calls in a row, the
second on a smaller problem, each equal to ccd_mbpt2 bit for bit (same
kernels, same summation order; the scratch array keeps the previous call's
partials, and every partial is rewritten before it is read).  Device memory
for the result and the caller's stream come from torch (plumbing only)."""
import numpy as np
import pytest

from coupled_cluster_doubles_amd import FP32, FP64, Batch, CCDError, Context, Model

pytestmark = pytest.mark.gpu


def test_noalloc_mbpt2_equals_mbpt2_on_two_calls_larger_then_smaller(require_gpu):
    import torch
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    big = Context(Model.electron_gas(1.0, 32, 4), FP64, matrix_free=True)    # 6,053 tiles
    small = Context(Model.pairing(8, 8, 0.5, 1.0), FP32)
    ref_big, ref_small = big.mbpt2(), small.mbpt2()
    torch.cuda.synchronize()
    big.mbpt2_noalloc(out.data_ptr())                       # the context's stream
    s = torch.cuda.Stream()
    small.mbpt2_noalloc(out.data_ptr() + 8, s.cuda_stream)  # a caller stream
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert got[0] == ref_big and got[1] == ref_small, (got, ref_big, ref_small)
    # the same again in the other order, both onto one caller stream
    out.fill_(float("nan"))
    torch.cuda.synchronize()
    small.mbpt2_noalloc(out.data_ptr() + 8, s.cuda_stream)
    big.mbpt2_noalloc(out.data_ptr(), s.cuda_stream)
    s.synchronize()
    got = out.cpu().numpy()
    assert got[0] == ref_big and got[1] == ref_small


def test_noalloc_mbpt2_empty_problem_and_batch_rejected(require_gpu):
    import torch
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    empty = Context(Model.electron_gas(1.0, 1, 1), FP64)    # no amplitudes
    torch.cuda.synchronize()
    empty.mbpt2_noalloc(out.data_ptr())
    torch.cuda.synchronize()
    assert out.item() == 0.0 == empty.mbpt2()
    b = Batch([Model.pairing(4, 4, 0.5, 1.0), Model.pairing(6, 6, 0.3, 1.0)], FP64)
    with pytest.raises(CCDError):
        b.mbpt2_noalloc(out.data_ptr())
    assert np.isfinite(b.mbpt2()).all()
