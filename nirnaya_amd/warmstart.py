"""Learned warm start for parametric LP families.

A family keeps A fixed and moves the right-hand side and the costs with a
parameter vector theta in R^64:
    b(theta) = b0 * (1 + 0.1 tanh(B theta)),   c(theta) = c0 + 0.05 C theta
(amplitudes 0.1 and 0.05 by default),
with fixed seeded N(0,1) matrices B [m][64] and C [n][64] and A, b0, c0 the
equality LP of nir_gen_equality_lp.

The warm start is a proper orthogonal decomposition (POD) of solution
snapshots - mean-centred, rank <= 256 for x and for y, computed from the
snapshot Gram matrix on the device (method of snapshots) - and an MLP
64 -> 512 -> 512 -> 512 (GELU, bf16 MFMA with FP32 accumulation, trained on
the device by the engine's own forward/backward kernels) that predicts the
POD coefficients of (x, y) from theta. Training data are solves the PDHG
engine itself produced. Every warm-started solve must reach the same KKT
tolerance as a cold solve; if it does not, the cold solve is the fallback and
its result is returned.

The reference has no learned component; this module has no counterpart there.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .engine import LP, MODE_PDLP, Engine, default_params, gen_equality_lp

THETA_DIM = 64


class ParametricFamily:
    """Fixed-A family of equality LPs (see the module docstring)."""

    def __init__(self, seed: int, m: int, n: int, per_col: int = 5, dim: int = THETA_DIM, b_amp: float = 0.1,
                 c_amp: float = 0.05):
        self.seed, self.m, self.n, self.dim = seed, m, n, dim
        self.b_amp, self.c_amp = b_amp, c_amp
        self.base = gen_equality_lp(seed, m, n, per_col)
        rng = np.random.default_rng(seed)
        self.B = rng.standard_normal((m, dim))
        self.C = rng.standard_normal((n, dim))
        self.b0 = self.base.rl.copy()
        self.c0 = self.base.c.copy()

    def thetas(self, count: int, seed: int) -> np.ndarray:
        return np.random.default_rng(seed).standard_normal((count, self.dim))

    def rhs_cost(self, theta: np.ndarray):
        """(b(theta), c(theta)) of the member at theta."""
        b = self.b0 * (1.0 + self.b_amp * np.tanh(self.B @ theta))
        c = self.c0 + self.c_amp * (self.C @ theta)
        return b, c

    def instance(self, theta: np.ndarray) -> LP:
        b, c = self.rhs_cost(theta)
        base = self.base
        return LP(self.m, self.n, base.start, base.index, base.value, np.ascontiguousarray(c), base.l, base.u,
                  np.ascontiguousarray(b), b.copy())


def top_eigh(G: np.ndarray, r: int, oversample: int = 32, sweeps: int = 8, seed: int = 0):
    """Leading r eigenpairs of the symmetric positive semidefinite G (N x N),
    descending, by block subspace iteration with a Rayleigh-Ritz step: Q spans
    G^q Omega (Omega N x (r + oversample) Gaussian), then the eigenpairs of
    Q^T G Q give Ritz values and vectors V = Q U, for which V^T G V is
    diagonal, so POD modes X V S^-1 are orthonormal. The leading pairs
    converge like (lambda_{r+oversample+1} / lambda_i)^sweeps; costs
    O(N^2 (r + oversample)) per sweep instead of a full O(N^3)
    eigendecomposition."""
    n = G.shape[0]
    k = min(n, r + oversample)
    Q = np.random.default_rng(seed).standard_normal((n, k))
    Q, _ = np.linalg.qr(G @ Q)
    for _ in range(sweeps):
        Q, _ = np.linalg.qr(G @ Q)
    w, U = np.linalg.eigh(Q.T @ G @ Q)
    order = np.argsort(w)[::-1][:r]
    return w[order], Q @ U[:, order]


class Pod:
    """Mean-centred POD of snapshots S [N][len]: modes [len][r] with orthonormal
    columns, singular values, and the coefficients of the training snapshots.
    The eigenpairs of the snapshot Gram matrix are exact (a full symmetric
    eigendecomposition) up to N = exact_limit snapshots and from block
    subspace iteration (top_eigh) beyond, where only the leading `rank` of N
    pairs are used. The modes and the mean live on the device
    (`engine.pod_basis`); decoding writes start points there, where the
    family engine reads them, and `modes` is a host copy made on demand."""

    def __init__(self, engine: Engine, S: np.ndarray, rank: int, rel_floor: float = 1e-12,
                 exact_limit: int = 2048, log=None):
        S = np.asarray(S, dtype=np.float64)
        self.mean = S.mean(axis=0)
        # Centred snapshots transposed to [len][N], a block of snapshots at a
        # time (the only extra copy of S).
        X = np.empty((S.shape[1], S.shape[0]))
        for i0 in range(0, S.shape[0], 256):
            X[:, i0:i0 + 256] = (S[i0:i0 + 256] - self.mean).T
        G = engine.pod_gram(X)                                  # X^T X on the device
        if log:
            log(f"POD: Gram matrix {G.shape[0]} x {G.shape[0]} of snapshots of length {X.shape[0]}")
        if G.shape[0] > exact_limit and 4 * rank < G.shape[0]:
            lam, V = top_eigh(G, rank)
        else:
            lam, V = np.linalg.eigh(G)
            order = np.argsort(lam)[::-1]
            lam, V = lam[order], V[:, order]
        keep = int(np.sum(lam > rel_floor * max(lam[0], 1e-300))) if lam.size else 0
        r = max(1, min(rank, keep))
        self.sing = np.sqrt(np.maximum(lam[:r], 0.0))
        self.basis = engine.pod_basis(X, V[:, :r] / self.sing, self.mean)   # modes X V S^-1, kept on the device
        self.train_coef = V[:, :r] * self.sing                   # [N][r] = modes^T (s - mean)
        self.rank = r
        self._modes = None

    @property
    def modes(self) -> np.ndarray:
        if self._modes is None:
            self._modes = self.basis.modes()
        return self._modes

    def encode(self, s: np.ndarray) -> np.ndarray:
        """Coefficients modes^T (s - mean) of one snapshot or of rows [count][len]."""
        s = np.asarray(s, dtype=np.float64)
        coef = self.basis.encode(s)
        return coef[0] if s.ndim == 1 else coef

    def decode(self, alpha: np.ndarray, to_host: bool = True):
        """mean + modes alpha for one coefficient vector or rows [count][r],
        computed on the device, where the rows stay for the family engine."""
        alpha = np.asarray(alpha, dtype=np.float64)
        rows = self.basis.decode(alpha, to_host)
        return rows[0] if (to_host and alpha.ndim == 1) else rows


def _pad64(k: int) -> int:
    return (k + 63) // 64 * 64


class WarmStart:
    """theta -> (x0, y0) through the MLP and the two POD bases."""

    def __init__(self, engine: Engine, pod_x: Pod, pod_y: Pod, hidden: int = 512, seed: int = 0,
                 max_rows: int = 512, dim: int = THETA_DIM, primal_weight: float = 0.0, out_width: int = 0):
        self.engine, self.pod_x, self.pod_y, self.dim = engine, pod_x, pod_y, dim
        self.max_rows = max_rows
        # The primal weight the family's training solves ended with (median),
        # carried into warm solves with the predicted point; 0 = engine default.
        self.primal_weight = primal_weight
        self.n_coef = pod_x.rank + pod_y.rank
        # Output width: the coefficients padded to a multiple of 64, or a
        # fixed width (512 for ranks 256 + 256) whose unused outputs train to 0.
        self.n_out = max(_pad64(self.n_coef), _pad64(out_width))
        self.net = engine.mlp(dim, hidden, self.n_out, max_rows, seed)
        self.mu = np.zeros(self.n_out)
        self.sd = np.ones(self.n_out)

    def fit(self, thetas: np.ndarray, epochs: int = 200, batch: int = 256, lr: float = 1e-3, seed: int = 0):
        """Train on the POD coefficients of the training snapshots, centred and
        divided by one scale for the x block and one for the y block: the
        squared error then weighs each mode by its share of the snapshot
        variance (the POD modes are orthonormal, so it is the squared error of
        the reconstructed x and y up to those two scales) instead of giving
        the noisy trailing modes the weight of the leading ones."""
        alpha = np.concatenate([self.pod_x.train_coef, self.pod_y.train_coef], axis=1)
        self.mu[:self.n_coef] = alpha.mean(axis=0)
        rx = self.pod_x.rank
        cx, cy = alpha[:, :rx] - self.mu[:rx], alpha[:, rx:] - self.mu[rx:self.n_coef]
        self.sd[:rx] = max(float(np.sqrt(np.mean(np.sum(cx * cx, axis=1)) / rx)), 1e-12)
        self.sd[rx:self.n_coef] = max(float(np.sqrt(np.mean(np.sum(cy * cy, axis=1)) / max(1, cy.shape[1]))), 1e-12)
        target = np.zeros((alpha.shape[0], self.n_out), dtype=np.float32)
        target[:, :self.n_coef] = (alpha - self.mu[:self.n_coef]) / self.sd[:self.n_coef]
        return self.net.train(thetas.astype(np.float32), target, epochs, batch, lr, seed)

    def coefficients(self, thetas: np.ndarray) -> np.ndarray:
        """POD coefficients [count][rank_x + rank_y] the network predicts for
        thetas [count][dim] (forward passes of at most max_rows rows)."""
        T = np.asarray(thetas, dtype=np.float32).reshape(-1, self.dim)
        out = np.concatenate([self.net.forward(T[i:i + self.max_rows]) for i in range(0, len(T), self.max_rows)])
        return (out.astype(np.float64) * self.sd + self.mu)[:, :self.n_coef]

    def predict(self, theta: np.ndarray):
        alpha = self.coefficients(theta)[0]
        rx = self.pod_x.rank
        return self.pod_x.decode(alpha[:rx]), self.pod_y.decode(alpha[rx:])

    def predict_batch(self, thetas: np.ndarray, to_host: bool = False):
        """Start points for thetas [count][dim], decoded on the device, where
        they stay as the pods' decode rows (the family engine's device starts,
        `family_solve(..., device_starts=ws)`); host copies (X0 [count][n],
        Y0 [count][m]) when to_host, else (None, None)."""
        alpha = self.coefficients(thetas)
        rx = self.pod_x.rank
        return self.pod_x.decode(alpha[:, :rx], to_host), self.pod_y.decode(alpha[:, rx:], to_host)


def family_solve(engine: Engine, family: ParametricFamily, thetas: np.ndarray, params, starts=None,
                 primal_weight: float = 0.0, lanes: int = 0, on_result=None, with_stats: bool = False,
                 device_starts: "WarmStart" = None):
    """Members theta_i solved by the family engine (nir_family_*: one copy of
    A, 64 members per wave, the single-LP engine's results bit for bit).
    starts[i] = (x0, y0) warm-starts member i with `primal_weight`; with
    device_starts (a WarmStart whose predict_batch ran on these thetas) member
    i starts from row i of its pods' decode rows, read on the device. Returns
    (results in theta order, or None with on_result, and seconds; plus the
    engine's nir_family_stats with `with_stats`)."""
    import time
    count = len(thetas)
    lanes = lanes or min(1024, max(64, (count + 63) // 64 * 64))
    F = engine.family(family.base, params, lanes=lanes)
    if device_starts is not None:
        px, py = device_starts.pod_x.basis, device_starts.pod_y.basis
        if px.rows < count or py.rows < count:
            raise ValueError("device starts: predict_batch did not decode a row for every member")
        F.set_device_starts(px, py)

    def member(i):
        b, c = family.rhs_cost(thetas[i])
        if device_starts is not None:
            return c, b, b, None, None, primal_weight
        if starts is None:
            return c, b, b
        x0, y0 = starts[i]
        return c, b, b, x0, y0, primal_weight
    try:
        t0 = time.perf_counter()
        res = F.solve(count, member, on_result)
        dt = time.perf_counter() - t0
        return (res, dt, F.stats()) if with_stats else (res, dt)
    finally:
        F.close()


@dataclass
class SolveRecord:
    status: int
    iterations: int
    seconds: float
    fallback: bool = False


@dataclass
class PipelineReport:
    n_train: int
    n_solved: int
    rank_x: int
    rank_y: int
    train_loss: List[float]
    cold: List[SolveRecord] = field(default_factory=list)
    warm: List[SolveRecord] = field(default_factory=list)
    train_solve_seconds: float = 0.0
    pod_seconds: float = 0.0
    train_seconds: float = 0.0
    cold_seconds: float = 0.0
    warm_seconds: float = 0.0
    predict_seconds: float = 0.0


def warm_params(params, primal_weight: float):
    """A copy of `params` with the initial primal weight of a warm solve."""
    from .engine import NirParams
    p = NirParams.from_buffer_copy(params)
    if primal_weight > 0:
        p.primal_weight = primal_weight
    return p


def solve_warm(engine: Engine, lp: LP, params, x0, y0, primal_weight: float = 0.0):
    """Warm-started solve; if it misses the tolerance, the cold solve (the
    caller's parameters) runs and its result is returned."""
    import time
    t0 = time.perf_counter()
    r = engine.solve(lp, warm_params(params, primal_weight), x0=x0, y0=y0)
    dt = time.perf_counter() - t0
    if r.status == 0:
        return r, SolveRecord(r.status, r.iterations, dt)
    t1 = time.perf_counter()
    c = engine.solve(lp, params)
    return c, SolveRecord(c.status, r.iterations + c.iterations, dt + time.perf_counter() - t1, fallback=True)


def run_pipeline(engine: Engine, family: ParametricFamily, n_train: int, n_test: int, rank: int = 256,
                 epochs: int = 200, batch: int = 256, lr: float = 1e-3, tol: float = 1e-6, precision: int = 0,
                 max_iters: int = 400000, seed: int = 7, log=None, train_tol: float = 0.0,
                 out_width: int = 512) -> PipelineReport:
    """Training solves by the family engine (to `train_tol`, default `tol`),
    POD bases and the MLP fitted on the device, then the held-out thetas
    solved cold and warm-started (the family engine again, to `tol`). A warm
    solve that misses the tolerance falls back to its cold solve."""
    import time
    say = log or (lambda *_: None)
    if precision != 0:
        raise ValueError("the family engine runs FP64")
    params = default_params(mode=MODE_PDLP, max_iters=max_iters, tol=tol)
    tparams = default_params(mode=MODE_PDLP, max_iters=max_iters, tol=train_tol or tol)
    th_train = family.thetas(n_train, seed)
    th_test = family.thetas(n_test, seed + 1)
    res, t_solve = family_solve(engine, family, th_train, tparams)
    ok = [i for i, r in enumerate(res) if r.status == 0]   # screen out failed / infeasible thetas
    say(f"training solves: {len(ok)}/{n_train} reached tolerance in {t_solve:.1f}s")
    t0 = time.perf_counter()
    pod_x = Pod(engine, np.stack([res[i].x for i in ok]), rank)
    pod_y = Pod(engine, np.stack([res[i].y for i in ok]), rank)
    t_pod = time.perf_counter() - t0
    omega = float(np.median([res[i].info.primal_weight for i in ok]))
    del res
    ws = WarmStart(engine, pod_x, pod_y, max_rows=max(64, batch), primal_weight=omega, out_width=out_width)
    t0 = time.perf_counter()
    losses = ws.fit(th_train[ok], epochs=epochs, batch=batch, lr=lr, seed=seed)
    t_fit = time.perf_counter() - t0
    say(f"POD ranks x {pod_x.rank} y {pod_y.rank} ({t_pod:.1f}s); MLP loss {losses[0]:.3e} -> {losses[-1]:.3e} "
        f"({t_fit:.1f}s)")
    rep = PipelineReport(n_train, len(ok), pod_x.rank, pod_y.rank, [float(v) for v in losses],
                         train_solve_seconds=t_solve, pod_seconds=t_pod, train_seconds=t_fit)
    cold, t_cold = family_solve(engine, family, th_test, params)
    t0 = time.perf_counter()
    ws.predict_batch(th_test)   # start points decoded on the device, read there by the family engine
    t_pred = time.perf_counter() - t0
    warm, t_warm = family_solve(engine, family, th_test, params, primal_weight=ws.primal_weight, device_starts=ws)
    per_c, per_w = t_cold / max(1, n_test), (t_warm + t_pred) / max(1, n_test)
    for c, w in zip(cold, warm):
        rep.cold.append(SolveRecord(c.status, c.iterations, per_c))
        if w.status == 0:
            rep.warm.append(SolveRecord(w.status, w.iterations, per_w))
        else:
            rep.warm.append(SolveRecord(c.status, w.iterations + c.iterations, per_w + per_c, fallback=True))
    rep.cold_seconds, rep.warm_seconds, rep.predict_seconds = t_cold, t_warm, t_pred
    ws.net.close()
    return rep
