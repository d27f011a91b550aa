"""CPU reference of the emulator MLP.  TEST INFRASTRUCTURE ONLY: used by
tests/ as the checker of the GPU MLP kernels (fg_mlp.hip), never on the
product path.

Same network as fg_mlp_create: dense layers widths[0] -> ... -> widths[L],
GELU(x) = x Phi(x) on the hidden layers, MSE loss mean((y - t)^2), flat
parameter vector [W_0 (out x in, row-major), b_0, W_1, b_1, ...].

With emulate_bf16=True the operands the GPU rounds to BF16 are rounded the
same way here (inputs, weights, hidden activations, the gradients fed to
the backward GEMMs) while every sum is carried in FP64; with False it is
the plain FP64 network, whose analytic gradient the tests pin against
central finite differences.

Parity pins: tests/test_mlp_oracle.py (finite differences, closed forms
of a one-layer network).
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, u + 0x7FFF + ((u >> 16) & 1), u) >> 16
    return ((r & 0xFFFF).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(params, widths):
    out, off = [], 0
    for l in range(len(widths) - 1):
        n_in, n_out = widths[l], widths[l + 1]
        W = np.asarray(params[off:off + n_in * n_out], dtype=np.float64).reshape(n_out, n_in)
        off += n_in * n_out
        b = np.asarray(params[off:off + n_out], dtype=np.float64)
        off += n_out
        out.append((W, b))
    return out


def forward(params, widths, x, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, widths)
    H = [r(x)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = H[-1] @ r(W).T + b
        Z.append(z)
        if l < len(layers) - 1:
            H.append(r(gelu(z)))
    return Z[-1], (H, Z, layers)


def loss_and_grad(params, widths, x, t, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (H, Z, layers) = forward(params, widths, x, emulate_bf16)
    n = y.size
    d = y - t
    loss = float((d * d).sum() / n)
    dz = 2.0 * d / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dzr = r(dz)
        gW = dzr.T @ H[l]
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dz = (dzr @ r(W)) * gelu_d(Z[l - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
