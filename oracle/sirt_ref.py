"""ORACLE: SIRT recurrence in float64 -- test infrastructure only.

The recurrence of sys_mat/sirt_dd.py:37-76 and tests/Test_SIRT/sirt_test.py:
21-40 evaluated with an explicit system matrix A (dense or scipy sparse):

    r = (b - A x) / W_sino   where W_sino != 0   (else A x)
    u = A^T r / W_img        where W_img  != 0   (else A^T r)
    x = x + gamma * u,       W_sino = A 1, W_img = A^T 1.

Pinned in tests by its fixed point: for consistent data b = A x_true with a
full-rank square-ish A the iterates converge to x_true.
"""
import numpy as np


def csr_matrix(ind, ln, off, nvox):
    import scipy.sparse as sp
    off = np.asarray(off, np.int64)
    return sp.csr_matrix((np.asarray(ln, np.float64), np.asarray(ind, np.int64), off),
                         shape=(len(off) - 1, nvox))


def sirt(A, b, iters, gamma=1.0, x0=None, W_sino=None, W_img=None, return_all=False):
    b = np.asarray(b, np.float64).reshape(-1)
    n = A.shape[1]
    x = np.zeros(n) if x0 is None else np.asarray(x0, np.float64).reshape(-1).copy()
    if W_sino is None:
        W_sino = A @ np.ones(n)
    if W_img is None:
        W_img = A.T @ np.ones(A.shape[0])
    W_sino = np.asarray(W_sino, np.float64).reshape(-1)
    W_img = np.asarray(W_img, np.float64).reshape(-1)
    m1, m2 = W_sino != 0, W_img != 0
    out = []
    for _ in range(iters):
        r = A @ x
        r[m1] = (b[m1] - r[m1]) / W_sino[m1]
        u = A.T @ r
        u[m2] /= W_img[m2]
        x = x + gamma * u
        if return_all:
            out.append(x.copy())
    return np.array(out) if return_all else x
