"""The allocation-free MLP backward (aq_mlp_backward_noalloc: deltas in a
static device array) against aq_mlp_backward with a caller workspace, through
the C ABI: two calls in a row, the second with fewer samples (the static
array keeps the first call's deltas beyond the second call's rows)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _backward_pair(L, params, x, t, wd):
    """(loss, grad) from aq_mlp_backward and from aq_mlp_backward_noalloc."""
    from autoquantizer_amd import _lib
    n = x.shape[0]
    H, OUT, NP = _lib.AQ_MLP_HIDDEN, _lib.AQ_MLP_OUT, _lib.AQ_MLP_PARAMS
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    act = torch.empty((n, 3 * H), dtype=torch.float32, device="cuda")
    y = torch.empty((n, OUT), dtype=torch.float32, device="cuda")
    _lib.check(L.aq_mlp_forward(_ptr(params), _ptr(x), n, _ptr(act), _ptr(y), st))
    delta = torch.empty((n, 3 * H + OUT), dtype=torch.float32, device="cuda")
    g_ws = torch.empty(NP, dtype=torch.float32, device="cuda")
    l_ws = torch.zeros(1, dtype=torch.float64, device="cuda")
    _lib.check(L.aq_mlp_backward(_ptr(params), _ptr(x), _ptr(act), _ptr(y), _ptr(t), n, wd, _ptr(delta),
                                 _ptr(g_ws), _ptr(l_ws), st))
    g_na = torch.full((NP,), float("nan"), dtype=torch.float32, device="cuda")
    l_na = torch.zeros(1, dtype=torch.float64, device="cuda")
    _lib.check(L.aq_mlp_backward_noalloc(_ptr(params), _ptr(x), _ptr(act), _ptr(y), _ptr(t), n, wd, _ptr(g_na),
                                         _ptr(l_na), st))
    torch.cuda.synchronize()
    return (l_ws.item(), g_ws.cpu().numpy()), (l_na.item(), g_na.cpu().numpy())


def test_noalloc_backward_equals_workspace_backward():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from autoquantizer_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(11)
    params = torch.as_tensor(rng.standard_normal(_lib.AQ_MLP_PARAMS) * 0.1, dtype=torch.float32, device="cuda")
    for n, wd in ((_lib.AQ_MLP_NOALLOC_MAX_SAMPLES, 3e-3), (301, 0.0)):
        print(f"noalloc backward: {n} samples, weight decay {wd}", flush=True)
        x = torch.as_tensor(rng.standard_normal((n, _lib.AQ_MLP_IN)), dtype=torch.float32, device="cuda")
        t = torch.as_tensor(rng.standard_normal((n, _lib.AQ_MLP_OUT)), dtype=torch.float32, device="cuda")
        (lw, gw), (ln, gn) = _backward_pair(L, params, x, t, wd)
        # the same kernels in the same order on the same inputs: identical gradients; the loss is
        # summed with atomics across workgroups, whose order varies, so it agrees to rounding
        assert np.isfinite(gn).all()
        np.testing.assert_array_equal(gn, gw)
        assert ln == pytest.approx(lw, rel=1e-12)

