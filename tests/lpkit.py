"""Small LPs used across the tests, with where each comes from."""
import os

import numpy as np

from nirnaya_amd.engine import LP

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INF = float("inf")


def pdhg_unit_lp() -> LP:
    """tests/unit_tests.cpp:316-331: min 2x1+3x2, x in [0,20]^2,
    x1+x2 in [4,20], x1+3x2 in [6,30]. Optimum x=(3,1), obj 9, y=(1.5,0.5)."""
    return LP.from_dense([[1.0, 1.0], [1.0, 3.0]], [2.0, 3.0], [0.0, 0.0], [20.0, 20.0], [4.0, 6.0], [20.0, 30.0])


def transport_lp() -> LP:
    """tools/bench_runner.cpp:23-39: 5x5 assignment LP, cost ((i+1)(j+2) mod 7)+1,
    rows and columns summing to 1. Optimum 13 (brute force over 5! permutations)."""
    N = 5
    A = np.zeros((2 * N, N * N))
    c = np.zeros(N * N)
    for i in range(N):
        for j in range(N):
            c[i * N + j] = ((i + 1) * (j + 2)) % 7 + 1
    for i in range(N):
        for j in range(N):
            A[2 * i, i * N + j] = 1.0      # row-sum rows come first per i
            A[2 * i + 1, j * N + i] = 1.0  # then the column-sum row for i
    ones = np.ones(2 * N)
    return LP.from_dense(A, c, np.zeros(N * N), np.ones(N * N), ones, ones)


def transport_optimum() -> float:
    import itertools
    N = 5
    best = INF
    for perm in itertools.permutations(range(N)):
        best = min(best, sum(((i + 1) * (perm[i] + 2)) % 7 + 1 for i in range(N)))
    return float(best)


def _dense(lp):
    A = np.zeros((lp.m, lp.n))
    for j in range(lp.n):
        A[lp.index[lp.start[j]:lp.start[j + 1]], j] = lp.value[lp.start[j]:lp.start[j + 1]]
    return A


def infeasible_lps():
    """LPs without an optimum, infeasible or unbounded by construction, with
    the kind of certificate that proves it: 'dual' (a Farkas ray y: the LP is
    infeasible) or 'primal' (a ray x of decreasing cost: no dual solution)."""
    from nirnaya_amd.engine import gen_equality_lp
    out = {}
    # x1 + x2 >= 5 and x1 + x2 <= 3 on [0, 10]^2: y = (1, -1) gives 5 - 3 > 0.
    out["two_rows_contradict"] = ("dual", LP.from_dense([[1.0, 1.0], [1.0, 1.0]], [1.0, 1.0], [0.0, 0.0],
                                                        [10.0, 10.0], [5.0, -INF], [INF, 3.0]))
    # min -x1, x1 - x2 <= 1, x >= 0: x = (1, 1) keeps feasibility and lowers the cost.
    out["ray_of_descent"] = ("primal", LP.from_dense([[1.0, -1.0]], [-1.0, 0.0], [0.0, 0.0], [INF, INF], [-INF],
                                                     [1.0]))
    # an equality LP with its first row repeated at right-hand side + 1
    base = gen_equality_lp(5, 200, 500, 5)
    A = _dense(base)
    out["repeated_row_shifted"] = ("dual", LP.from_dense(np.vstack([A, A[:1]]), base.c, base.l, base.u,
                                                         np.append(base.rl, base.rl[0] + 1.0),
                                                         np.append(base.ru, base.ru[0] + 1.0)))
    # an equality LP with a column and its negative, both unbounded above, the
    # first with cost -1: their difference is a free direction of descent
    base = gen_equality_lp(6, 200, 500, 5)
    A = _dense(base)
    out["free_direction"] = ("primal", LP.from_dense(np.hstack([A, A[:, :1], -A[:, :1]]),
                                                     np.append(base.c, [-1.0, 0.0]), np.append(base.l, [0.0, 0.0]),
                                                     np.append(base.u, [INF, INF]), base.rl, base.ru))
    return out


def dense_line_lp(seed, m, n, row_len, col_len, per_col=3):
    """Equality LP (feasible by construction, x in [0, 10]) with row 1 holding
    `row_len` and column 2 holding `col_len` nonzeros beside `per_col` random
    ones per column."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    ri = rng.integers(0, m, per_col * n)
    ci = np.repeat(np.arange(n), per_col)
    ri = np.concatenate([ri, np.full(row_len, 1), rng.choice(m, col_len, replace=False)])
    ci = np.concatenate([ci, rng.choice(n, row_len, replace=False), np.full(col_len, 2)])
    A = sp.csc_matrix((rng.standard_normal(ri.size), (ri, ci)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    b = A @ rng.uniform(0, 1, n)
    c = A.T @ rng.standard_normal(m) + rng.uniform(0, 1, n)
    return LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64), c,
              np.zeros(n), np.full(n, 10.0), b.copy(), b.copy())


def _csc(lp):
    import scipy.sparse as sp
    return sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n))


def _from_csc(A, c, l, u, rl, ru):
    A = A.tocsc()
    A.sort_indices()
    return LP(A.shape[0], A.shape[1], A.indptr.astype(np.int64), A.indices.astype(np.int32),
              A.data.astype(np.float64), np.asarray(c, float), np.asarray(l, float), np.asarray(u, float),
              np.asarray(rl, float), np.asarray(ru, float))


def long_line_infeasible_lps():
    """As infeasible_lps, on LPs whose rays run through a row and a column of
    more than 4,096 nonzeros (both certificates' products then take the long
    lines' summation order)."""
    import scipy.sparse as sp
    out = {}
    # The dense row (row 1) repeated at right-hand side + 1: y on the pair.
    base = dense_line_lp(31, 5000, 6000, 5000, 4500)
    A = _csc(base)
    out["dense_row_repeated_shifted"] = ("dual", _from_csc(
        sp.vstack([A, A[1]]), base.c, base.l, base.u, np.append(base.rl, base.rl[1] + 1.0),
        np.append(base.ru, base.ru[1] + 1.0)))
    # The dense column (column 2) and its negative appended, unbounded above,
    # the first with cost -1: their difference is a free direction of descent.
    base = dense_line_lp(32, 4300, 4600, 4500, 4200, per_col=2)
    A = _csc(base)
    out["dense_column_free_direction"] = ("primal", _from_csc(
        sp.hstack([A, A[:, 2], -A[:, 2]]), np.append(base.c, [-1.0, 0.0]), np.append(base.l, [0.0, 0.0]),
        np.append(base.u, [INF, INF]), base.rl, base.ru))
    return out


def read_mps_lp(path):
    """Minimal free-MPS reader for the fixtures (N/L/G/E rows, COLUMNS, RHS,
    UP/LO bounds, OBJSENSE MAX). Returns (LP with the file's own costs,
    sense (+1 min / -1 max), column names)."""
    rows, kinds, cols, entries, cost, rhs, lo, up = [], {}, [], {}, {}, {}, {}, {}
    sense, obj, section = 1, None, None
    with open(path) as f:
        for line in f:
            if not line.strip() or line.startswith("*"):
                continue
            tok = line.split()
            if not line[0].isspace():
                section = tok[0].upper()
                continue
            if section == "OBJSENSE":
                sense = -1 if tok[0].upper().startswith("MAX") else 1
            elif section == "ROWS":
                if tok[0].upper() == "N":
                    obj = tok[1]
                else:
                    rows.append(tok[1])
                    kinds[tok[1]] = tok[0].upper()
            elif section == "COLUMNS":
                cn = tok[0]
                if cn not in cost:
                    cols.append(cn)
                    cost[cn] = 0.0
                for k in range(1, len(tok) - 1, 2):
                    if tok[k] == obj:
                        cost[cn] += float(tok[k + 1])
                    else:
                        entries[(tok[k], cn)] = float(tok[k + 1])
            elif section == "RHS":
                for k in range(len(tok) % 2, len(tok) - 1, 2):
                    rhs[tok[k]] = float(tok[k + 1])
            elif section == "BOUNDS":
                b, cn, v = tok[0].upper(), tok[2], float(tok[3])
                (up if b == "UP" else lo)[cn] = v
    m, n = len(rows), len(cols)
    A = np.zeros((m, n))
    for (rn, cn), v in entries.items():
        A[rows.index(rn), cols.index(cn)] = v
    rl = np.array([rhs.get(r, 0.0) if kinds[r] in ("G", "E") else -INF for r in rows])
    ru = np.array([rhs.get(r, 0.0) if kinds[r] in ("L", "E") else INF for r in rows])
    c = np.array([cost[cn] for cn in cols])
    lp = LP.from_dense(A, c, [lo.get(cn, 0.0) for cn in cols], [up.get(cn, INF) for cn in cols], rl, ru)
    return lp, sense, cols


def read_sol(path):
    obj, vals = None, {}
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "#":
                if tok[1] == "objective":
                    obj = float(tok[2])
                continue
            vals[tok[0]] = float(tok[1])
    return obj, vals


CRUDE_MPS = os.path.join(GOLDEN, "MRPL_Crude_Blending_LP.mps")
CRUDE_SOL = os.path.join(GOLDEN, "MRPL_Crude_Blending_LP.sol")


def random_long_line_lp(seed):
    """Seeded LP of random shape with 1-3 rows and 0-2 columns of 4,097 to
    ~12,000 nonzeros beside 1-4 random nonzeros per column, and mixed bound
    types: free, one-sided and boxed columns; equality, one-sided, ranged
    and free rows. Feasible (rows hold A x0 for x0 inside the bounds)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(7000 + seed)
    m = int(rng.integers(4200, 7000))
    n = int(rng.integers(4500, 13000))
    per = rng.integers(1, 5, n)
    ri = [rng.integers(0, m, int(per.sum()))]
    ci = [np.repeat(np.arange(n), per)]
    for r in rng.choice(m, int(rng.integers(1, 4)), replace=False):
        k = int(rng.integers(4097, min(n, 12000) + 1))
        ri.append(np.full(k, r))
        ci.append(rng.choice(n, k, replace=False))
    for c in rng.choice(n, int(rng.integers(0, 3)), replace=False):
        k = int(rng.integers(4097, m + 1))
        ri.append(rng.choice(m, k, replace=False))
        ci.append(np.full(k, c))
    ri, ci = np.concatenate(ri), np.concatenate(ci)
    A = sp.csc_matrix((rng.standard_normal(ri.size), (ri, ci)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    kind = rng.integers(0, 4, n)   # 0 boxed, 1 lower only, 2 upper only, 3 free
    l = np.where((kind == 0) | (kind == 1), 0.0, -INF)
    u = np.where(kind == 0, rng.uniform(1, 10, n), np.where(kind == 2, 10.0, INF))
    x0 = np.clip(rng.uniform(-1, 5, n), np.where(np.isfinite(l), l, -1), np.where(np.isfinite(u), u, 5))
    b = A @ x0
    rk = rng.integers(0, 4, m)     # 0 equality, 1 >=, 2 <=, 3 ranged
    rl = np.where(rk == 2, -INF, b - np.where(rk == 0, 0.0, rng.uniform(0, 1, m)))
    ru = np.where(rk == 1, INF, b + np.where(rk == 0, 0.0, rng.uniform(0, 1, m)))
    c = A.T @ rng.standard_normal(m) + np.where(kind == 3, 0.0, rng.uniform(0, 1, n))
    return LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64), c, l, u,
              rl, ru)
