"""Closed-form LASSO optima near the zero threshold.

For min 1/2||Ax-b||^2 + lam||x||_1 with g = A^T b: x = 0 is optimal iff
lam >= ||g||_inf.  Just below, with j = argmax |g_j|, the one-sparse point
x_j = sign(g_j)(|g_j| - lam)/||a_j||^2 is optimal as long as every other
column's correlation with the residual stays within lam (checked here)."""
import numpy as np
import scipy.sparse as sp


def closed_form(A, b, lam):
    A = sp.csc_matrix(A)
    g = A.T @ b
    x = np.zeros(A.shape[1])
    if lam >= np.max(np.abs(g)):
        return x, 0.5 * float(b @ b)
    j = int(np.argmax(np.abs(g)))
    aj = A[:, j].toarray().ravel()
    x[j] = np.sign(g[j]) * (abs(g[j]) - lam) / float(aj @ aj)
    r = b - aj * x[j]
    corr = np.abs(A.T @ r)
    corr[j] = 0.0
    assert corr.max() <= lam, "one-sparse point is not optimal for this lam"
    return x, 0.5 * float(r @ r) + lam * abs(x[j])
