"""CPU checks of the warm-start oracle (oracle/mlp_oracle.py) and of the POD
host logic (nirnaya_amd.warmstart.Pod) with numpy stand-ins for the two device
products.

Pins: the bf16 rounding against hand-computed values (ties to even); the
back-propagation of the oracle against central finite differences of its own
loss in plain FP64 (a dropped term, a wrong sign or a transposed operand
fails); GELU's derivative against finite differences; the POD against
numpy's SVD of the centred snapshot matrix (singular values, subspace,
reconstruction)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_bf16_round_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 3.0e38, 1e-30],
                 dtype=np.float32)
    r = mo.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins.
    # 1 + 3 2^-9 is above the midpoint: rounds up to 1 + 2^-7.
    # 1 + 2^-7 + 2^-8 ties between 1 + 2^-7 (odd) and 1 + 2^-6 (even): up.
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == np.float32(1.0 + 2 ** -6)
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e38) / 3.0e38 < 2 ** -8 and abs(r[6] - 1e-30) / 1e-30 < 2 ** -8
    # every result has its low 16 bits clear
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-6, 6, 97)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(z))) < 1e-8
    # closed forms: GELU(0) = 0, GELU'(0) = 1/2, GELU(z) -> z for large z
    assert mo.gelu(np.array([0.0]))[0] == 0.0 and abs(mo.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    assert abs(mo.gelu(np.array([10.0]))[0] - 10.0) < 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_backprop_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    n_in, hid, n_out, rows = 3, 5, 4, 7
    count = sum(int(np.prod(s)) for s in mo.layout(n_in, hid, n_out))
    p = rng.standard_normal(count) * 0.7
    X = rng.standard_normal((rows, n_in))
    T = rng.standard_normal((rows, n_out))
    loss, g = mo.loss_and_grad(p, X, T, n_in, hid, n_out, emulate_bf16=False)
    fd = np.zeros(count)
    h = 1e-6
    for i in range(count):
        e = np.zeros(count)
        e[i] = h
        lp, _ = mo.loss_and_grad(p + e, X, T, n_in, hid, n_out, emulate_bf16=False)
        lm, _ = mo.loss_and_grad(p - e, X, T, n_in, hid, n_out, emulate_bf16=False)
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    # forward agrees with the loss's own forward
    out = mo.forward(p, X, n_in, hid, n_out, emulate_bf16=False)
    assert abs(np.mean((out - T) ** 2) - loss) < 1e-14


def test_emulated_rounding_changes_little():
    rng = np.random.default_rng(3)
    n_in, hid, n_out = 64, 64, 64
    count = sum(int(np.prod(s)) for s in mo.layout(n_in, hid, n_out))
    p = rng.standard_normal(count) * 0.1
    X = rng.standard_normal((32, n_in))
    T = rng.standard_normal((32, n_out))
    _, g64 = mo.loss_and_grad(p, X, T, n_in, hid, n_out, emulate_bf16=False)
    _, g16 = mo.loss_and_grad(p, X, T, n_in, hid, n_out, emulate_bf16=True)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert 1e-5 < rel < 5e-2   # bf16 rounding is visible but small


class _NumpyBasis:
    """numpy stand-in for the device-resident basis (modes X V, the mean)."""

    def __init__(self, X, V, mean):
        self._m, self._mean, self.rank, self.rows = X @ V, mean, V.shape[1], 0

    def modes(self):
        return self._m.copy()

    def decode(self, alpha, to_host=True):
        alpha = np.atleast_2d(alpha)
        self.rows = alpha.shape[0]
        return self._mean + alpha @ self._m.T if to_host else None

    def encode(self, S):
        return (np.atleast_2d(S) - self._mean) @ self._m


class _NumpyPodEngine:
    """numpy stand-ins for the device products of the POD."""

    @staticmethod
    def pod_gram(X):
        return X.T @ X

    @staticmethod
    def pod_basis(X, V, mean):
        return _NumpyBasis(X, V, mean)


def test_pod_host_logic_matches_svd():
    from nirnaya_amd.warmstart import Pod
    rng = np.random.default_rng(11)
    N, L, r_true = 40, 300, 6
    S = rng.standard_normal((N, r_true)) @ rng.standard_normal((r_true, L)) * 3.0 + rng.standard_normal(L)
    S += 1e-6 * rng.standard_normal((N, L))
    pod = Pod(_NumpyPodEngine(), S, rank=10)
    Xc = S - S.mean(axis=0)
    U, sv, Vt = np.linalg.svd(Xc.T, full_matrices=False)   # columns of U: spatial modes
    k = pod.rank
    assert np.allclose(pod.sing, sv[:k], rtol=1e-9, atol=1e-9 * sv[0])
    # orthonormal modes spanning the same leading subspace
    assert np.allclose(pod.modes.T @ pod.modes, np.eye(k), atol=1e-9)
    P1 = pod.modes[:, :r_true] @ pod.modes[:, :r_true].T
    P2 = U[:, :r_true] @ U[:, :r_true].T
    assert np.linalg.norm(P1 - P2) < 1e-8
    # coefficients of the training snapshots and reconstruction
    assert np.allclose(pod.train_coef, np.stack([pod.encode(s) for s in S]), atol=1e-8 * sv[0])
    rec = np.stack([pod.decode(pod.encode(s)) for s in S])
    assert np.max(np.abs(rec - S)) < 1e-4


def test_pod_subspace_iteration_matches_svd_on_leading_modes():
    # Above exact_limit snapshots the POD takes the leading eigenpairs of the
    # Gram matrix by block subspace iteration: against numpy's SVD, the
    # leading singular values and subspace agree, the modes are orthonormal,
    # and the training coefficients are the modes' projections.
    from nirnaya_amd.warmstart import Pod
    rng = np.random.default_rng(12)
    N, L, r = 600, 900, 32
    basis = np.linalg.qr(rng.standard_normal((L, 120)))[0]
    coef = rng.standard_normal((N, 120)) * (np.arange(1, 121) ** -1.5) * 10.0
    S = coef @ basis.T + 1e-4 * rng.standard_normal((N, L))
    pod = Pod(_NumpyPodEngine(), S, rank=r, exact_limit=100)
    Xc = S - S.mean(axis=0)
    U, sv, _ = np.linalg.svd(Xc.T, full_matrices=False)
    assert pod.rank == r
    assert np.allclose(pod.sing[:16], sv[:16], rtol=1e-9)
    assert np.allclose(pod.sing, sv[:r], rtol=1e-3)
    assert np.allclose(pod.modes.T @ pod.modes, np.eye(r), atol=1e-9)
    P1 = pod.modes[:, :8] @ pod.modes[:, :8].T
    P2 = U[:, :8] @ U[:, :8].T
    assert np.linalg.norm(P1 - P2) < 1e-8
    assert np.allclose(pod.train_coef, np.stack([pod.encode(s) for s in S]), atol=1e-8 * sv[0])
