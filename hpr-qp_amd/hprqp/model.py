"""Model building: the host half of HPRQP.jl's API.

Mirrors (same names, argument meaning and errors):
  qp_formulation ........ src/utils.jl:67-188
  build_from_mps ........ src/utils.jl:889-925
  build_from_QAbc ....... src/utils.jl:926-1003
  build_from_mat ........ src/utils.jl:1030-1090
  build_from_ABST ....... src/utils.jl:1133-1180
  build_from_Ab_lambda .. src/utils.jl:1225-1267
  Q operator types ...... src/Q_operators/{LASSO,QAP}_operator.jl (CPU structs)
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import Optional, Union

import numpy as np
import scipy.sparse as sp

from .mps import read_mps


# --------------------------------------------------------------------------
# Q operators (AbstractQOperatorCPU, src/Q_operators/Q_operator_interface.jl)
# --------------------------------------------------------------------------
class AbstractQOperator:
    """Plugin interface (src/Q_operators/Q_operator_interface.jl:5-92):
    get_problem_size / get_temp_size / name, the CPU apply Qmap, and the
    device apply Qmap_device that the solver calls.

    A user operator implements Qmap_device(x_ptr, Qx_ptr, stream): enqueue
    Qx = Q x on the HIP stream `stream` for device arrays of n doubles at the
    integer addresses x_ptr and Qx_ptr (the GPU Qmap! of the reference), for
    example by calling its own HIP library.  Such an operator is registered
    with hprqp_qop_create_user through ctypes callbacks.  The built-in
    operators below are created natively instead."""

    name = "USER"

    def get_problem_size(self) -> int:
        raise NotImplementedError

    def get_temp_size(self) -> int:
        return 0

    def Qmap(self, x: np.ndarray) -> np.ndarray:
        """Reference-semantics CPU apply (used only to describe the operator)."""
        raise NotImplementedError

    def Qmap_device(self, x_ptr: int, Qx_ptr: int, stream: int) -> None:
        """Device apply on `stream`; a user operator must provide it."""
        raise NotImplementedError


class LASSO_Q_operator(AbstractQOperator):
    """Q = A^T A for a LASSO data matrix A (src/Q_operators/LASSO_operator.jl:50-65)."""

    name = "LASSO"

    def __init__(self, A, b=None):
        self.A = sp.csr_matrix(A, dtype=np.float64)
        self.A.sort_indices()
        self.b = None if b is None else np.asarray(b, dtype=np.float64)

    def get_problem_size(self):
        return self.A.shape[1]

    def get_temp_size(self):
        return self.A.shape[0]

    def Qmap(self, x):
        return self.A.T @ (self.A @ x)


class QAP_Q_operator(AbstractQOperator):
    """Q(X) = 2(A X B - S X - X T), X = reshape(x, n, n) column-major
    (src/Q_operators/QAP_operator.jl:7-30)."""

    name = "QAP"

    def __init__(self, A, B, S, T):
        self.A, self.B, self.S, self.T = (np.asarray(M, dtype=np.float64) for M in (A, B, S, T))
        self.n = self.A.shape[0]
        for M in (self.A, self.B, self.S, self.T):
            if M.shape != (self.n, self.n):
                raise AssertionError("A, B, S, T must be square and of the same size")

    def get_problem_size(self):
        return self.n * self.n

    def get_temp_size(self):
        return self.n * self.n

    def Qmap(self, x):
        X = x.reshape(self.n, self.n, order="F")
        QX = 2.0 * (self.A @ X @ self.B - self.S @ X - X @ self.T)
        return QX.reshape(-1, order="F")


class DiagonalPlusLowRank_Q_operator(AbstractQOperator):
    """Q = Diag(diag) + U U^T, U n x k: the custom operator of the reference's
    operator guide (docs/src/guide/q_operators_overview.md:121-139)."""

    name = "DiagonalPlusLowRank"

    def __init__(self, diag, U):
        self.diag = np.asarray(diag, dtype=np.float64)
        self.U = np.asarray(U, dtype=np.float64)
        if self.U.ndim != 2 or self.U.shape[0] != self.diag.size:
            raise AssertionError("U must be n x k with n = length(diag)")

    def get_problem_size(self):
        return self.diag.size

    def get_temp_size(self):
        return self.U.shape[1]

    def Qmap(self, x):
        return self.diag * x + self.U @ (self.U.T @ x)


QType = Union[sp.spmatrix, AbstractQOperator]


@dataclass
class QP_info_cpu:
    """The formulated model (src/structs.jl:98-164)."""

    Q: QType
    c: np.ndarray
    A: sp.csc_matrix
    AL: np.ndarray
    AU: np.ndarray
    l: np.ndarray
    u: np.ndarray
    obj_constant: float = 0.0
    lam: float = 0.0
    deleted_rows: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))

    @property
    def n(self):
        return len(self.c)

    @property
    def m(self):
        return self.A.shape[0]

    @property
    def AT(self):
        return self.A.T.tocsc()


def _as_csc(M):
    if isinstance(M, np.ndarray):
        return sp.csc_matrix(M)
    return sp.csc_matrix(M, dtype=np.float64)


def qp_formulation(Q, c, A, AL, AU, l, u, c0: float = 0.0, verbose: bool = True) -> QP_info_cpu:
    """Validate and preprocess (src/utils.jl:67-188)."""
    Q = _as_csc(Q)
    A = _as_csc(A)
    c = np.asarray(c, dtype=np.float64).copy()
    AL = np.asarray(AL, dtype=np.float64).copy()
    AU = np.asarray(AU, dtype=np.float64).copy()
    l = np.asarray(l, dtype=np.float64).copy()
    u = np.asarray(u, dtype=np.float64).copy()
    mQ, nQ = Q.shape
    if mQ != nQ:
        raise ValueError(f"Q matrix must be square. Got size ({mQ}, {nQ}).")
    n = nQ
    if Q.nnz > 0:
        diff = (Q - Q.T).tocsc()
        dev = np.max(np.abs(diff.data)) if diff.nnz else 0.0
        if dev > 1e-10:
            warnings.warn(f"Q matrix is not symmetric (max deviation: {dev}). Symmetrizing Q = 0.5*(Q + Q').")
            Q = (0.5 * (Q + Q.T)).tocsc()
            Q.eliminate_zeros()
    if len(c) != n:
        raise ValueError(f"Dimension mismatch: Q is {n}×{n} but c has length {len(c)}.")
    if len(l) != n:
        raise ValueError(f"Dimension mismatch: Q is {n}×{n} but l has length {len(l)}.")
    if len(u) != n:
        raise ValueError(f"Dimension mismatch: Q is {n}×{n} but u has length {len(u)}.")
    mA, nA = A.shape
    if nA != n:
        raise ValueError(f"Dimension mismatch: Q is {n}×{n} but A has {nA} columns.")
    if len(AL) != mA:
        raise ValueError(f"Dimension mismatch: A has {mA} rows but AL has length {len(AL)}.")
    if len(AU) != mA:
        raise ValueError(f"Dimension mismatch: A has {mA} rows but AU has length {len(AU)}.")
    bad = np.nonzero(l > u)[0]
    if bad.size:
        raise ValueError(f"Infeasible variable bounds: l > u at indices: {list(bad[:5] + 1)}")
    bad = np.nonzero(AL > AU)[0]
    if bad.size:
        raise ValueError(f"Infeasible constraint bounds: AL > AU at rows: {list(bad[:5] + 1)}")
    if Q.nnz and not np.all(np.isfinite(Q.data)):
        raise ValueError("Q matrix contains NaN or Inf values.")
    if not np.all(np.isfinite(c)):
        raise ValueError("c vector contains NaN or Inf values.")
    if A.nnz and not np.all(np.isfinite(A.data)):
        raise ValueError("A matrix contains NaN or Inf values.")
    if np.any(np.isnan(AL)) or np.any(np.isnan(AU)):
        raise ValueError("Constraint bounds AL or AU contain NaN values.")
    if np.any(np.isnan(l)) or np.any(np.isnan(u)):
        raise ValueError("Variable bounds l or u contain NaN values.")
    # remove empty rows and free rows (src/utils.jl:153-163)
    rowsum = np.asarray(abs(A).sum(axis=1)).ravel() if mA else np.zeros(0)
    del_row = np.union1d(np.nonzero(rowsum == 0)[0], np.nonzero((AL == -np.inf) & (AU == np.inf))[0])
    if del_row.size:
        keep = np.setdiff1d(np.arange(mA), del_row)
        A = A[keep, :].tocsc()
        AL = AL[keep]
        AU = AU[keep]
        if verbose:
            print(f"Deleted {del_row.size} rows of A that are all zeros.")
    A.sort_indices()
    Q.sort_indices()
    return QP_info_cpu(Q, c, A, AL, AU, l, u, float(c0), 0.0, del_row.astype(np.int64))


def build_from_mps(filename: str, verbose: bool = True) -> QP_info_cpu:
    """src/utils.jl:889-925."""
    if not (filename.endswith(".mps") or filename.endswith(".MPS")):
        raise ValueError("Unsupported file format. Please provide a .mps file.")
    Q, c, A, lcon, ucon, lvar, uvar, c0 = read_mps(filename)
    return qp_formulation(Q, c, A, lcon, ucon, lvar, uvar, c0, verbose=verbose)


def build_from_QAbc(Q, c, A, AL, AU, l, u, obj_constant: float = 0.0, verbose: bool = True) -> QP_info_cpu:
    """src/utils.jl:926-1003: dense inputs are converted to sparse."""
    return qp_formulation(Q, c, A, AL, AU, l, u, obj_constant, verbose=verbose)


def build_from_ABST(A, B, S, T, verbose: bool = True) -> QP_info_cpu:
    """QAP relaxation with doubly stochastic constraints (src/utils.jl:1133-1180)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    op = QAP_Q_operator(A, B, S, T)
    ee = np.ones((1, n))
    I = sp.identity(n, format="csr")
    Acon = sp.vstack([sp.kron(ee, I), sp.kron(I, ee)]).tocsc()
    Acon.sort_indices()
    return QP_info_cpu(op, np.zeros(n * n), Acon, np.ones(2 * n), np.ones(2 * n), np.zeros(n * n),
                       np.full(n * n, np.inf), 0.0, 0.0)


def build_from_Ab_lambda(A, b, lam: float, verbose: bool = True) -> QP_info_cpu:
    """LASSO: min 1/2||Ax-b||^2 + lam ||x||_1 (src/utils.jl:1225-1267)."""
    A = sp.csr_matrix(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m, n = A.shape
    if len(b) != m:
        raise AssertionError("Length of b must match number of rows in A")
    if not lam > 0:
        raise AssertionError("Lambda must be positive")
    c = -(A.T @ b)
    obj_constant = 0.5 * float(np.dot(b, b))
    op = LASSO_Q_operator(A, b)
    return QP_info_cpu(op, np.asarray(c, dtype=np.float64), sp.csc_matrix((0, n)), np.zeros(0), np.zeros(0),
                       np.full(n, -np.inf), np.full(n, np.inf), obj_constant, float(lam))


def build_from_mat(filename: str, problem_type: str = "QAP", verbose: bool = True) -> QP_info_cpu:
    """QAP (A, B, S, T) or LASSO (A, b) from a MAT file (src/utils.jl:1030-1090)."""
    import scipy.io as sio

    data = sio.loadmat(filename)
    if problem_type == "QAP":
        return build_from_ABST(*(np.asarray(data[k], dtype=np.float64) for k in "ABST"), verbose=verbose)
    if problem_type == "LASSO":
        A = sp.csr_matrix(data["A"], dtype=np.float64)
        b = np.asarray(data["b"], dtype=np.float64).ravel()
        lam = 0.001 * np.max(np.abs(A.T @ b))
        if verbose:
            print("Auto-selected lambda = ", lam)
        return build_from_Ab_lambda(A, b, lam, verbose=verbose)
    raise ValueError(f"Unknown problem_type: {problem_type}. Supported types are 'QAP' and 'LASSO'.")
