"""vcl_mse_loss_backward_noalloc against vcl_mse_loss_backward through the C ABI.

The no-allocation variant keeps its block partials in the library's own device array instead
of a caller workspace; both run the same kernels in the same order, so the gradient and the
loss must be bit-identical.  Two calls in a row, the second smaller than the first, check that
the shared array's leftover contents from the first call do not leak into the second; a size
above the stated maximum is rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from vcl_amd import _lib as L
    lib = L.load()
    lib.vcl_mse_loss_backward_noalloc.argtypes = [C.POINTER(L.LossParams), C.c_void_p]
    lib.vcl_mse_loss_backward_noalloc.restype = C.c_int
    lib.vcl_mse_loss_noalloc_max_size.argtypes = []
    lib.vcl_mse_loss_noalloc_max_size.restype = C.c_size_t
    return L, lib


def _call(L, lib, fn, pred, target, with_workspace):
    n = pred.numel()
    grad = torch.empty_like(pred)
    loss = torch.full((1,), float("nan"), device="cuda")
    ws = torch.empty(int(lib.vcl_loss_workspace_size(n)), dtype=torch.uint8, device="cuda") if with_workspace else None
    p = L.LossParams(pred.data_ptr(), target.data_ptr(), grad.data_ptr(), ws.data_ptr() if ws is not None else None,
                     n, 1, 1, n, loss.data_ptr())
    st = fn(C.byref(p), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, st
    torch.cuda.synchronize()
    return grad.cpu().numpy(), float(loss.cpu()[0])


def test_noalloc_mse_matches_workspace_entry_point():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L, lib = _lib()
    rng = np.random.default_rng(7)
    for n in (3 * 2 ** 20 + 7, 4096):  # second call smaller: leftovers of the first must not count
        a = torch.from_numpy(rng.normal(size=n).astype(np.float32)).cuda()
        b = torch.from_numpy(rng.normal(size=n).astype(np.float32)).cuda()
        g_ref, l_ref = _call(L, lib, lib.vcl_mse_loss_backward, a, b, True)
        g_new, l_new = _call(L, lib, lib.vcl_mse_loss_backward_noalloc, a, b, False)
        print(f"n={n}: loss workspace {l_ref!r} no-alloc {l_new!r}", flush=True)
        assert np.array_equal(g_new, g_ref)
        assert l_new == l_ref
        ref = 0.5 * np.sum((a.cpu().numpy().astype(np.float64) - b.cpu().numpy()) ** 2)
        assert abs(l_new - ref) <= 1e-5 * ref
    # above the maximum: rejected on the host (VCL_ERR_OOM), nothing launched or read
    mx = int(lib.vcl_mse_loss_noalloc_max_size())
    assert mx == 2 ** 28
    dummy = torch.zeros(4, device="cuda")
    p = L.LossParams(dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), None, mx + 1, 1, 1, mx + 1, dummy.data_ptr())
    assert lib.vcl_mse_loss_backward_noalloc(C.byref(p), None) == 2
