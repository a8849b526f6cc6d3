"""fgpr_tv_static (the TV denoiser with its workspace in the library's static
device array, allocating nothing) against fgpr_tv with a caller workspace,
through the C ABI: bitwise equal on two calls in a row, the second on a
smaller model, and a model too large for the static workspace is rejected
with an error before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _tv_both(lib, data, lamda=0.03, rho=1.3, num_iter=200, tol=1e-6):
    M, N = data.shape
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    nbytes = lib.fgpr_tv_workspace_bytes(M, N)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=data.device)
    a = torch.full_like(data, np.nan)
    b = torch.full_like(data, np.nan)
    rc = lib.fgpr_tv(ctypes.c_void_p(data.data_ptr()), M, N, lamda, rho, num_iter, tol,
                     ctypes.c_void_p(ws.data_ptr()), nbytes, ctypes.c_void_p(a.data_ptr()), st)
    assert rc == 0, lib.fgpr_last_error()
    rc = lib.fgpr_tv_static(ctypes.c_void_p(data.data_ptr()), M, N, lamda, rho, num_iter, tol,
                            ctypes.c_void_p(b.data_ptr()), st)
    assert rc == 0, lib.fgpr_last_error()
    torch.cuda.synchronize()
    return a, b


def test_static_workspace_tv_equals_caller_workspace(hip_device):
    from fastgpr._native import lib
    rng = np.random.default_rng(21)
    for M, N in ((220, 120), (61, 37)):                 # the second call is smaller
        print("case", M, N, flush=True)
        data = torch.tensor(5 + rng.standard_normal((M, N)).astype(np.float32),
                            device=hip_device)
        a, b = _tv_both(lib, data)
        assert torch.isfinite(a).all() and not torch.equal(a, data)
        assert torch.equal(a, b), (M, N)
    big = (4000, 2000)
    assert lib.fgpr_tv_workspace_bytes(*big) > (64 << 20)
    out = torch.empty(1, device=hip_device)
    rc = lib.fgpr_tv_static(ctypes.c_void_p(out.data_ptr()), big[0], big[1], 0.03, 1.3, 10, 1e-6,
                            ctypes.c_void_p(out.data_ptr()),
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0 and b"static" in lib.fgpr_last_error()
