"""HIP kernels of the learned-correction network against the CPU oracle
(oracle/cnn.py, which emulates the bfloat16 storage of the GPU path).

Tolerances: the GPU accumulates in FP32 in its own order, so an activation
near a bfloat16 rounding boundary can land one bf16 ulp (2^-8 relative) away
from the oracle's; over three layers that gives relative L2 differences of a
few 1e-3.  The checks use 1e-2 and, as a control, require the GPU to be much
closer to the bf16-emulating oracle than the float64 network is."""
import ctypes

import numpy as np
import pytest

from oracle import cnn as C

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


class Net:
    def __init__(self, H, W, B):
        import torch
        from meeuuw_amd import _lib
        self.torch, self.lib = torch, _lib
        self.h = ctypes.c_void_p()
        _lib.call("mw_cnn_create", H, W, B, ctypes.byref(self.h))
        self.st = torch.cuda.current_stream().cuda_stream
        self.H, self.W = H, W

    def set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self.lib.call("mw_cnn_set_params", self.h, p.ctypes.data, self.st)

    def get(self, what):
        out = np.empty(C.NPARAM, dtype=np.float32)
        self.lib.call("mw_cnn_get_" + what, self.h, out.ctypes.data, self.st)
        return out

    def forward(self, X0):
        t = self.torch
        x = t.tensor(np.ascontiguousarray(X0, dtype=np.float32), device="cuda")
        c = t.empty(X0.shape[:3], device="cuda", dtype=t.float32)
        self.lib.call("mw_cnn_forward", self.h, X0.shape[0], x.data_ptr(), c.data_ptr(), self.st)
        self._keep = x
        return c.cpu().numpy()

    def backward(self, dC):
        t = self.torch
        d = t.tensor(np.ascontiguousarray(dC, dtype=np.float32), device="cuda")
        self.lib.call("mw_cnn_backward", self.h, dC.shape[0], d.data_ptr(), self.st)
        t.cuda.synchronize()
        return self.get("grads")

    def close(self):
        self.lib.call("mw_cnn_destroy", self.h)


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = C.init_params(seed)
    d = C.unpack(p.copy())
    for k in ("b1", "b2", "b3", "b4"):
        d[k] = (0.1 * rng.standard_normal(d[k].shape)).astype(np.float32)
    d["W4"] = d["W4"] * 10
    return C.pack(d).astype(np.float32)


BLOCKS = [("W1", 0, 1152), ("b1", 1152, 1184), ("W2", 1184, 10400), ("b2", 10400, 10432),
          ("W3", 10432, 19648), ("b3", 19648, 19680), ("W4", 19680, 19968), ("b4", 19968, 19969)]


@pytest.mark.parametrize("H,W,B", [(13, 37, 3), (8, 32, 1), (1, 5, 2)])
def test_forward_backward_vs_oracle(H, W, B):
    rng = np.random.default_rng(H * 100 + W)
    params = random_params(H + W)
    X0 = rng.standard_normal((B, H, W, 4)).astype(np.float32)
    dC = rng.standard_normal((B, H, W)).astype(np.float32)
    net = Net(H, W, B)
    try:
        net.set(params)
        np.testing.assert_array_equal(net.get("params"), params)
        c_gpu = net.forward(X0)
        g_gpu = net.backward(dC)
    finally:
        net.close()
    c_ref, cache = C.forward(params, X0)
    g_ref = C.backward(cache, dC)
    c_ex, cache_ex = C.forward(params, X0, exact=True)
    g_ex = C.backward(cache_ex, dC)
    print(H, W, B, "C rel", rel(c_gpu, c_ref), "exact vs bf16", rel(c_ex, c_ref))
    assert rel(c_gpu, c_ref) < 1e-2
    assert rel(c_gpu, c_ref) < 0.5 * rel(c_ex, c_ref) or rel(c_gpu, c_ref) < 1e-5
    for name, a, b in BLOCKS:
        r = rel(g_gpu[a:b], g_ref[a:b])
        print(name, r, rel(g_ex[a:b], g_ref[a:b]))
        assert r < 2e-2, name


def test_forward_full_grid():
    """The pressure grid of a 256x256 mesh (257 x 257 pixels), two samples."""
    H = W = 257
    rng = np.random.default_rng(7)
    params = random_params(11)
    X0 = rng.standard_normal((2, H, W, 4)).astype(np.float32)
    net = Net(H, W, 2)
    try:
        net.set(params)
        c_gpu = net.forward(X0)
    finally:
        net.close()
    c_ref, _ = C.forward(params, X0)
    print("257^2 C rel", rel(c_gpu, c_ref))
    assert rel(c_gpu, c_ref) < 1e-2


def test_adam_step_vs_oracle():
    H, W, B = 9, 11, 2
    rng = np.random.default_rng(3)
    params = random_params(5)
    X0 = rng.standard_normal((B, H, W, 4)).astype(np.float32)
    dC = rng.standard_normal((B, H, W)).astype(np.float32)
    net = Net(H, W, B)
    try:
        net.set(params)
        net.forward(X0)
        g = net.backward(dC)
        net.lib.call("mw_cnn_adam", net.h, 1e-3, net.st)
        p1 = net.get("params")
        net.forward(X0)
        g2 = net.backward(dC)
        net.lib.call("mw_cnn_adam", net.h, 1e-3, net.st)
        p2 = net.get("params")
    finally:
        net.close()
    q1, m, v = C.adam(params.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    np.testing.assert_allclose(p1, q1, rtol=1e-6, atol=1e-8)
    q2, m, v = C.adam(p1.astype(np.float64), g2.astype(np.float64), m, v, 2, 1e-3)
    np.testing.assert_allclose(p2, q2, rtol=1e-6, atol=1e-8)
