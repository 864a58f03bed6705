"""CPU oracle of the learned warm start's network: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as a checker; the product never calls it.

The warm-start MLP has no counterpart in the reference (the reference has no
learned component); this restates the device kernels' arithmetic in FP64 with
numpy:
    Z1 = X W1^T + b1, H1 = GELU(Z1), Z2 = H1 W2^T + b2, H2 = GELU(Z2),
    Z3 = H2 W3^T + b3, loss = mean((Z3 - T)^2),
and its gradient by back-propagation. With `emulate_bf16` the values the
device rounds to bf16 before a GEMM (inputs, weights, hidden activations,
back-propagated errors) are rounded here the same way (round to nearest even)
and every sum is taken in FP64, so the device result differs from this only
by its FP32 accumulation order. Pins: central finite differences of the loss
in plain FP64 (tests/test_mlp_oracle.py) fix the back-propagation formulas.
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover - scipy is part of the image
    _erf = np.vectorize(math.erf)

SQRT1_2 = 1.0 / math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def bf16_round(a):
    """Round float32 values to the nearest bf16 (ties to even); returns float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def layout(n_in, hidden, n_out):
    """Shapes of the flat parameter vector, in the device's order."""
    return [(hidden, n_in), (hidden,), (hidden, hidden), (hidden,), (n_out, hidden), (n_out,)]


def unpack(params, n_in, hidden, n_out):
    out, o = [], 0
    for shp in layout(n_in, hidden, n_out):
        k = int(np.prod(shp))
        out.append(np.asarray(params[o:o + k], dtype=np.float64).reshape(shp))
        o += k
    return out


def loss_and_grad(params, X, T, n_in, hidden, n_out, emulate_bf16=True):
    """Mean squared error and its gradient (flat, FP64)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in, hidden, n_out)
    if emulate_bf16:
        rb = lambda a: bf16_round(np.asarray(a, dtype=np.float32)).astype(np.float64)
        f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    else:
        rb = lambda a: np.asarray(a, dtype=np.float64)
        f32 = rb
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    rows = X.shape[0]
    Xb, W1b, W2b, W3b = rb(X), rb(W1), rb(W2), rb(W3)
    Z1 = f32(Xb @ W1b.T + b1)
    H1 = rb(gelu(Z1))
    Z2 = f32(H1 @ W2b.T + b2)
    H2 = rb(gelu(Z2))
    Z3 = f32(H2 @ W3b.T + b3)
    diff = Z3 - T
    loss = float(np.mean(diff * diff))
    D3 = f32(2.0 * diff / (rows * n_out))
    D3b = rb(D3)
    gW3 = D3b.T @ H2
    gb3 = D3.sum(0)
    D2 = f32(f32(D3b @ W3b) * gelu_grad(Z2))
    D2b = rb(D2)
    gW2 = D2b.T @ H1
    gb2 = D2.sum(0)
    D1 = f32(f32(D2b @ W2b) * gelu_grad(Z1))
    D1b = rb(D1)
    gW1 = D1b.T @ Xb
    gb1 = D1.sum(0)
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad


def forward(params, X, n_in, hidden, n_out, emulate_bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in, hidden, n_out)
    rb = (lambda a: bf16_round(np.asarray(a, dtype=np.float32)).astype(np.float64)) if emulate_bf16 else \
        (lambda a: np.asarray(a, dtype=np.float64))
    H1 = rb(gelu(rb(X) @ rb(W1).T + b1))
    H2 = rb(gelu(H1 @ rb(W2).T + b2))
    return H2 @ rb(W3).T + b3
