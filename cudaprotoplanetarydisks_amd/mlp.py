"""Python handle on the MLP kernels of libfargo_hip.so (fg_mlp_*)."""
import ctypes

import numpy as np

from .native import call, lib

_pf = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(_pf)


def bf16_gemm(A, B):
    """C = A B with A, B rounded to BF16, FP32 accumulation (on the GPU)."""
    A, B = _f32(A), _f32(B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    C = np.empty((M, N), np.float32)
    call("fg_bf16_gemm", _p(A), _p(B), M, N, K, _p(C))
    return C


class MLP:
    def __init__(self, widths, max_batch=256, seed=1, device=0):
        self.widths = list(widths)
        w = (ctypes.c_int * len(widths))(*widths)
        h = ctypes.c_void_p()
        call("fg_mlp_create", w, len(widths) - 1, max_batch, seed, device, ctypes.byref(h))
        self.h = h
        self.nparam = lib().fg_mlp_nparam(h)

    def close(self):
        if getattr(self, "h", None):
            lib().fg_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def params(self):
        out = np.empty(self.nparam, np.float32)
        call("fg_mlp_get_params", self.h, _p(out))
        return out.astype(np.float64)

    def set_params(self, p):
        p = _f32(p)
        call("fg_mlp_set_params", self.h, _p(p))

    def forward(self, x):
        x = _f32(x)
        y = np.empty((x.shape[0], self.widths[-1]), np.float32)
        call("fg_mlp_forward", self.h, _p(x), x.shape[0], _p(y))
        return y.astype(np.float64)

    def gradient(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(self.nparam, np.float32)
        loss = np.empty(1, np.float32)
        call("fg_mlp_gradient", self.h, _p(x), _p(t), x.shape[0], _p(g), _p(loss))
        return float(loss[0]), g.astype(np.float64)

    def train(self, x, t, batch, steps, lr):
        x, t = _f32(x), _f32(t)
        losses = np.empty(steps, np.float32)
        call("fg_mlp_train", self.h, _p(x), _p(t), x.shape[0], batch, steps, ctypes.c_float(lr), _p(losses))
        return losses.astype(np.float64)
