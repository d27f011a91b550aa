"""MLP kernels (BF16 MFMA, FP32 accumulation) against the CPU reference
MLP (oracle/mlp_oracle.py): the GEMM, the forward pass and the gradient of
every parameter, then a training run."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

from cudaprotoplanetarydisks_amd.mlp import MLP, bf16_gemm  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(64, 64, 32), (70, 45, 37), (256, 512, 256), (5, 256, 1), (256, 5, 256)])
def test_bf16_gemm_matches_fp64_of_rounded_operands(shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C = bf16_gemm(A, B)
    ref = mo.bf16(A) @ mo.bf16(B)
    err = np.abs(C - ref).max() / max(np.abs(ref).max(), 1e-30)
    assert err < 1e-5, err


def test_forward_matches_reference():
    widths = [5, 64, 64, 48]
    rng = np.random.default_rng(3)
    with MLP(widths, max_batch=96, seed=9) as net:
        x = rng.uniform(-1, 1, size=(96, 5)).astype(np.float32)
        y = net.forward(x)
        ref, _ = mo.forward(net.params(), widths, x, emulate_bf16=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 1e-2, err


def test_gradient_matches_reference():
    widths = [5, 64, 64, 48]
    rng = np.random.default_rng(4)
    with MLP(widths, max_batch=64, seed=11) as net:
        x = rng.uniform(-1, 1, size=(64, 5)).astype(np.float32)
        t = rng.normal(size=(64, 48)).astype(np.float32)
        loss, g = net.gradient(x, t)
        ref_loss, ref_g = mo.loss_and_grad(net.params(), widths, x, t, emulate_bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    rel = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    cos = float(g @ ref_g / np.linalg.norm(g) / np.linalg.norm(ref_g))
    assert rel < 1e-2 and cos > 0.9999, (rel, cos)


def test_training_fits_a_smooth_map():
    """5 -> 4 x 64 -> 32 fits y_k(x) = sin(x . a_k) on 256 samples."""
    widths = [5, 64, 64, 64, 64, 32]
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(256, 5)).astype(np.float32)
    a = rng.normal(size=(5, 32))
    t = np.sin(x @ a).astype(np.float32)
    with MLP(widths, max_batch=256, seed=1) as net:
        losses = net.train(x, t, batch=256, steps=1500, lr=2e-3)
        y = net.forward(x)
    assert losses[-1] < losses[0] / 50, (losses[0], losses[-1])
    assert np.sqrt(((y - t) ** 2).mean()) < 0.1


def test_bad_batch_is_rejected():
    from cudaprotoplanetarydisks_amd.native import FargoError
    with MLP([5, 16, 4], max_batch=8) as net:
        with pytest.raises(FargoError):
            net.forward(np.zeros((9, 5), np.float32))
