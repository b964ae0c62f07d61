"""Seeded synthetic instances of the benchmark workloads.

All generators use numpy's PCG64 (``default_rng(seed)``), so every machine
with the same numpy regenerates identical data.  Sizes are parameters so the
same construction serves small parity cases and the full workloads.

  random_sparse_qp  random sparse convex QP (n=10,000, m=5,000 at full size)
  grid_qp           structured grid QP (1,414 x 1,414 cells at full size)
  lasso             sparse LASSO with implicit Q = A^T A (3,308 x 150,360)
  qap               QAP relaxation, 128 x 128 (dense operator)
  portfolio         long-only factor-model portfolio QP (batch member)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .model import (DiagonalPlusLowRank_Q_operator, QAP_Q_operator, QP_info_cpu, build_from_ABST,
                    build_from_Ab_lambda, build_from_QAbc)


def _distinct_cols(rng, rows, per_row, n):
    """per_row distinct uniform column indices for each of `rows` rows."""
    out = np.empty((rows, per_row), dtype=np.int64)
    for i in range(rows):
        out[i] = rng.choice(n, size=per_row, replace=False)
    return out


def random_sparse_qp(n=10_000, m=5_000, seed=0, a_per_row=5, r_rows=None, r_per_col=4, eq_frac=0.3,
                     box=10.0, verbose=False) -> QP_info_cpu:
    """A: exactly a_per_row nnz per row at distinct uniform columns, N(0,1).
    Q = R^T R + diag(d), R is r_rows x n with r_per_col nnz per column N(0,1),
    d ~ U[0.01, 1], stored explicitly.  c ~ N(0,1).  x0 ~ U[-1,1], b = A x0.
    The first eq_frac of the rows are equalities; the rest get
    l = b - U[0,1], u = b + U[0,1].  Box x in [-box, box]."""
    rng = np.random.default_rng(seed)
    r_rows = r_rows if r_rows is not None else max(1, n // 10)
    cols = _distinct_cols(rng, m, a_per_row, n)
    vals = rng.standard_normal((m, a_per_row))
    A = sp.csr_matrix((vals.ravel(), (np.repeat(np.arange(m), a_per_row), cols.ravel())), shape=(m, n))
    rrows = np.empty((n, r_per_col), dtype=np.int64)
    for j in range(n):
        rrows[j] = rng.choice(r_rows, size=r_per_col, replace=False)
    rvals = rng.standard_normal((n, r_per_col))
    R = sp.csc_matrix((rvals.ravel(), (rrows.ravel(), np.repeat(np.arange(n), r_per_col))), shape=(r_rows, n))
    d = rng.uniform(0.01, 1.0, n)
    Q = (R.T @ R + sp.diags(d)).tocsc()
    Q = (0.5 * (Q + Q.T)).tocsc()
    c = rng.standard_normal(n)
    x0 = rng.uniform(-1.0, 1.0, n)
    b = A @ x0
    meq = int(round(eq_frac * m))
    lo = b.copy()
    hi = b.copy()
    lo[meq:] -= rng.uniform(0.0, 1.0, m - meq)
    hi[meq:] += rng.uniform(0.0, 1.0, m - meq)
    return build_from_QAbc(Q, c, A, lo, hi, np.full(n, -box), np.full(n, box), verbose=verbose)


def grid_qp(N=1414, seed=0, verbose=False) -> QP_info_cpu:
    """N x N grid of cells (n = N^2), cell (i, j) -> variable i*N + j.

    Q = 5-point Laplacian (Dirichlet, diagonal 4, off-diagonal -1) + 1e-3 I.
    A: one row per horizontal cell pair {(i,2k), (i,2k+1)} (m = N*floor(N/2)):
    +1 on the pair's two cells and -1 on its six neighbours (i,2k-1),
    (i,2k+2), (i-1,2k), (i-1,2k+1), (i+1,2k), (i+1,2k+1) that exist, i.e.
    5-8 nonzeros per row, entries +-1.  x0 ~ U[0,1], AL = AU = A x0,
    x in [0, inf), c ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    n = N * N
    idx = np.arange(n).reshape(N, N)
    # Laplacian
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [np.full(n, 4.0 + 1e-3)]
    for di, dj in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        si = slice(max(0, -di), N - max(0, di))
        sj = slice(max(0, -dj), N - max(0, dj))
        ti = slice(max(0, di), N - max(0, -di))
        tj = slice(max(0, dj), N - max(0, -dj))
        rows.append(idx[si, sj].ravel())
        cols.append(idx[ti, tj].ravel())
        vals.append(np.full(rows[-1].size, -1.0))
    Q = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    # edge operator
    K = N // 2
    ii, kk = np.meshgrid(np.arange(N), np.arange(K), indexing="ij")
    ii = ii.ravel()
    kk = kk.ravel()
    m = ii.size
    r_all, c_all, v_all = [], [], []
    row = np.arange(m)
    for (di, dj, s) in ((0, 0, 1.0), (0, 1, 1.0), (0, -1, -1.0), (0, 2, -1.0), (-1, 0, -1.0), (-1, 1, -1.0),
                        (1, 0, -1.0), (1, 1, -1.0)):
        ci = ii + di
        cj = 2 * kk + dj
        ok = (ci >= 0) & (ci < N) & (cj >= 0) & (cj < N)
        r_all.append(row[ok])
        c_all.append(ci[ok] * N + cj[ok])
        v_all.append(np.full(ok.sum(), s))
    A = sp.csr_matrix((np.concatenate(v_all), (np.concatenate(r_all), np.concatenate(c_all))), shape=(m, n))
    x0 = rng.uniform(0.0, 1.0, n)
    b = A @ x0
    c = rng.standard_normal(n)
    return build_from_QAbc(Q, c, A, b, b.copy(), np.zeros(n), np.full(n, np.inf), verbose=verbose)


def lasso(m=3308, n=150_360, mean_nnz=1400.0, seed=0, lam_factor=1e-3, verbose=False) -> QP_info_cpu:
    """Row nnz ~ log-normal with the given mean (sigma=1), distinct uniform
    columns, values log-normal(0,1), each row scaled to unit l2; b ~ N(0,1);
    lambda = lam_factor * ||A^T b||_inf; phi = lambda ||x||_1."""
    rng = np.random.default_rng(seed)
    mu = np.log(mean_nnz) - 0.5
    k = np.clip(np.rint(rng.lognormal(mu, 1.0, m)).astype(np.int64), 1, n)
    indptr = np.concatenate([[0], np.cumsum(k)])
    indices = np.empty(indptr[-1], dtype=np.int64)
    for i in range(m):
        indices[indptr[i]:indptr[i + 1]] = np.sort(rng.choice(n, size=k[i], replace=False))
    data = rng.lognormal(0.0, 1.0, indptr[-1])
    for i in range(m):
        s = slice(indptr[i], indptr[i + 1])
        data[s] /= np.linalg.norm(data[s])
    A = sp.csr_matrix((data, indices, indptr), shape=(m, n))
    b = rng.standard_normal(m)
    lam = lam_factor * np.max(np.abs(A.T @ b))
    return build_from_Ab_lambda(A, b, lam, verbose=verbose)


def qap(nq=128, seed=0, density=0.1, grid=32, verbose=False):
    """Flow F: symmetric, given density, integers U{1..10}; distance D:
    Manhattan distances of nq points U{0..grid-1}^2.  Q(X) = F X D + s X + X t
    with s = t = max(0, -lambda_min(D (x) F))/2, so Q is PSD.  Returned as the
    reference's QAP model with A = F/2, B = D, S = -s/2 I, T = -t/2 I."""
    rng = np.random.default_rng(seed)
    upper = np.triu(rng.random((nq, nq)) < density, 1)
    F = np.where(upper, rng.integers(1, 11, (nq, nq)), 0).astype(np.float64)
    F = F + F.T
    pts = rng.integers(0, grid, (nq, 2))
    D = np.abs(pts[:, None, 0] - pts[None, :, 0]) + np.abs(pts[:, None, 1] - pts[None, :, 1])
    D = D.astype(np.float64)
    ef = np.linalg.eigvalsh(F)
    ed = np.linalg.eigvalsh(D)
    lam_min = min(ef.min() * ed.min(), ef.min() * ed.max(), ef.max() * ed.min(), ef.max() * ed.max())
    shift = max(0.0, -lam_min) / 2.0
    I = np.eye(nq)
    model = build_from_ABST(F / 2.0, D, -shift / 2.0 * I, -shift / 2.0 * I, verbose=verbose)
    return model, dict(F=F, D=D, shift=shift)


def portfolio(n, k, seed=0, explicit=True):
    """min 1/2 x^T (F F^T + D) x - mu^T x  s.t. 1^T x = 1, 0 <= x <= 0.1.
    F: n x k N(0, 0.1^2); D ~ U[0.01, 0.1]; mu ~ N(0.05, 0.02^2).
    explicit=True stores Q = F F^T + D as a sparse matrix (scaled like any
    sparse QP); explicit=False keeps it as the diagonal-plus-low-rank
    operator (unscaled, as every operator Q).  Returns (model, factors)."""
    rng = np.random.default_rng(seed)
    F = rng.normal(0.0, 0.1, (n, k))
    Dg = rng.uniform(0.01, 0.1, n)
    mu = rng.normal(0.05, 0.02, n)
    A = sp.csr_matrix(np.ones((1, n)))
    if explicit:
        Q = sp.csc_matrix(F @ F.T + np.diag(Dg))
        model = build_from_QAbc(Q, -mu, A, np.ones(1), np.ones(1), np.zeros(n), np.full(n, 0.1), verbose=False)
    else:
        model = QP_info_cpu(DiagonalPlusLowRank_Q_operator(Dg, F), -mu, sp.csc_matrix(A), np.ones(1), np.ones(1),
                            np.zeros(n), np.full(n, 0.1))
    return model, dict(F=F, D=Dg, mu=mu)


def portfolio_batch(count=4096, seed=0, n_range=(200, 5000), k_range=(10, 50)):
    """The batch of long-only factor-model portfolio QPs: instance i has
    n ~ U{n_range}, k ~ U{k_range} drawn from seed, data from seed + 1 + i."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, count)
    ks = rng.integers(k_range[0], k_range[1] + 1, count)
    return [portfolio(int(ns[i]), int(ks[i]), seed=seed + 1 + i, explicit=False) for i in range(count)]
