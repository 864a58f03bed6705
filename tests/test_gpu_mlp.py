"""GPU checks of the warm-start kernels: the bf16 MFMA GEMM in all four operand
layouts against numpy on the same bf16-rounded inputs, the network's forward
and gradient against the CPU oracle (oracle/mlp_oracle.py, which rounds to
bf16 at the same points and sums in FP64), training, and the POD products.

Tolerances: the GEMM sums bf16 products in FP32, so it agrees with the FP64
sum of the same products to ~1e-6 relative per element (bound 1e-5 * sqrt(K)
of the row scale). Forward and gradient can also differ where an FP32 value
sits next to a bf16 rounding boundary; the bound is 2e-3 relative in the
Frobenius norm of each parameter block and 2e-2 of the block's largest entry
elementwise."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_layouts_match_numpy(engine, ta, tb):
    rng = np.random.default_rng(5)
    M, N, K = 128, 192, 96
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((K, N) if tb else (N, K)).astype(np.float32)
    C = engine.gemm_check(A, B, ta, tb)
    Ab = mo.bf16_round(A).astype(np.float64)
    Bb = mo.bf16_round(B).astype(np.float64)
    ref = (Ab.T if ta else Ab) @ (Bb if tb else Bb.T)
    scale = np.sqrt(K) * np.max(np.abs(Ab)) * np.max(np.abs(Bb))
    err = np.max(np.abs(C - ref))
    print(f"layout ta={ta} tb={tb}: max err {err:.2e} (scale {scale:.1f})")
    assert err <= 1e-5 * scale


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_large_tile_gemm_layouts_match_numpy(engine, ta, tb):
    # 2048 x 2048 outputs = 256 tiles of 128 x 128 and K a multiple of 64:
    # the large-tile kernel.
    rng = np.random.default_rng(6)
    M, N, K = 2048, 2048, 192
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((K, N) if tb else (N, K)).astype(np.float32)
    C = engine.gemm_check(A, B, ta, tb)
    Ab = mo.bf16_round(A).astype(np.float64)
    Bb = mo.bf16_round(B).astype(np.float64)
    ref = (Ab.T if ta else Ab) @ (Bb if tb else Bb.T)
    scale = np.sqrt(K) * np.max(np.abs(Ab)) * np.max(np.abs(Bb))
    err = np.max(np.abs(C - ref))
    print(f"large tiles ta={ta} tb={tb}: max err {err:.2e} (scale {scale:.1f})")
    assert err <= 1e-5 * scale


def _net(engine, seed=0):
    n_in, hid, n_out = 64, 128, 192
    net = engine.mlp(n_in, hid, n_out, max_rows=256, seed=seed)
    return net, n_in, hid, n_out


def test_forward_and_gradient_match_oracle(engine):
    net, n_in, hid, n_out = _net(engine, 3)
    rng = np.random.default_rng(8)
    p = net.params()
    X = rng.standard_normal((100, n_in)).astype(np.float32)   # 100 rows: padded to 128 on the device
    T = rng.standard_normal((100, n_out)).astype(np.float32)
    out = net.forward(X)
    ref = mo.forward(p, X, n_in, hid, n_out, emulate_bf16=True)
    assert np.max(np.abs(out - ref)) <= 2e-2 * np.max(np.abs(ref))
    assert np.linalg.norm(out - ref) <= 2e-3 * np.linalg.norm(ref)
    loss, g = net.gradient(X, T)
    lref, gref = mo.loss_and_grad(p, X, T, n_in, hid, n_out, emulate_bf16=True)
    print(f"loss {loss:.6e} oracle {lref:.6e}")
    assert abs(loss - lref) <= 1e-4 * lref
    o = 0
    for shp in mo.layout(n_in, hid, n_out):
        k = int(np.prod(shp))
        gb, rb = g[o:o + k], gref[o:o + k]
        rel = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        print(f"block {shp}: rel {rel:.2e} max {np.max(np.abs(gb - rb)):.2e} / {np.max(np.abs(rb)):.2e}")
        assert rel <= 2e-3
        assert np.max(np.abs(gb - rb)) <= 2e-2 * np.max(np.abs(rb))
        o += k
    net.close()


def test_training_fits_a_smooth_map(engine):
    net, n_in, hid, n_out = _net(engine, 4)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((1024, n_in)).astype(np.float32)
    W = rng.standard_normal((n_in, n_out)) / np.sqrt(n_in)
    T = np.tanh(X @ W).astype(np.float32)
    losses = net.train(X, T, epochs=40, batch=128, lr=2e-3, seed=1)
    print("epoch losses", losses[0], losses[-1])
    assert losses[-1] < 0.2 * losses[0]
    rows, ms, flops = net.train_stats()
    assert rows >= 40 * 1024 and ms > 0 and flops > 0
    net.close()


def test_pod_products_match_numpy(engine):
    rng = np.random.default_rng(10)
    X = rng.standard_normal((1000, 37))
    G = engine.pod_gram(X)
    assert np.allclose(G, X.T @ X, rtol=1e-12, atol=1e-9)
    V = rng.standard_normal((37, 5))
    assert np.allclose(engine.pod_project(X, V), X @ V, rtol=1e-12, atol=1e-10)


def test_device_pod_basis_matches_numpy(engine):
    # The basis kept on the device: modes = X V, decode = mean + alpha modes^T
    # (one DGEMM onto the broadcast mean), encode = (s - mean) modes; the
    # decode rows stay on the device (rows) for the family engine.
    rng = np.random.default_rng(12)
    X = rng.standard_normal((1500, 40))
    V = rng.standard_normal((40, 7))
    mean = rng.standard_normal(1500)
    pod = engine.pod_basis(X, V, mean)
    try:
        M = X @ V
        assert np.allclose(pod.modes(), M, rtol=1e-12, atol=1e-10)
        alpha = rng.standard_normal((33, 7))
        rows = pod.decode(alpha)
        assert pod.rows == 33
        assert np.allclose(rows, mean + alpha @ M.T, rtol=1e-12, atol=1e-9)
        S = rng.standard_normal((9, 1500))
        assert np.allclose(pod.encode(S), (S - mean) @ M, rtol=1e-12, atol=1e-9)
        assert pod.decode(alpha[:5], to_host=False) is None and pod.rows == 5
    finally:
        pod.close()
