"""ORACLE: SIRT recurrence in float64 -- test infrastructure only.

The recurrence of sys_mat/sirt_dd.py:37-76 and tests/Test_SIRT/sirt_test.py:
21-40 evaluated with an explicit system matrix A (dense or scipy sparse):

    r = (b - A x) / W_sino   where W_sino != 0   (else A x)
    u = A^T r / W_img        where W_img  != 0   (else A^T r)
    x = x + gamma * u,       W_sino = A 1, W_img = A^T 1.

The two reference drivers differ in start and count:
  sirt_dd_convention:   x0 = 1e-5, iters - 1 updates   (sirt_dd.py:46-51)
  sirt_test_convention: x0 = 0,    iters updates, every iterate returned
                        (sirt_test.py:23-35; slot i-1 of the zeroed
                        image_iters is the start at i = 0)

Pinned (tests/test_sirt_oracle.py) by the weighted least-squares fixed point
(numpy.linalg.lstsq of R^1/2 A x = R^1/2 b, R = diag(1/W_sino)) on an
inconsistent overdetermined system, a hand-worked step with a zero row and
a zero column, and planted mistakes (`mutate`) that those pins reject.
"""
import numpy as np


def csr_matrix(ind, ln, off, nvox):
    import scipy.sparse as sp
    off = np.asarray(off, np.int64)
    return sp.csr_matrix((np.asarray(ln, np.float64), np.asarray(ind, np.int64), off),
                         shape=(len(off) - 1, nvox))


MUTATIONS = ("inverted_mask", "extra_update", "no_image_weight", "residual_sign")


def sirt(A, b, iters, gamma=1.0, x0=None, W_sino=None, W_img=None, return_all=False,
         mutate=None):
    """`iters` updates from x0 (default 0).  mutate plants one of MUTATIONS
    (for the tests that show the pins reject it)."""
    if mutate is not None and mutate not in MUTATIONS:
        raise ValueError(mutate)
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
    if mutate == "inverted_mask":
        m1, m2 = ~m1, ~m2
    out = []
    steps = iters + 1 if mutate == "extra_update" else iters
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(steps):
            r = A @ x
            if mutate == "residual_sign":
                r[m1] = (r[m1] - b[m1]) / W_sino[m1]
            else:
                r[m1] = (b[m1] - r[m1]) / W_sino[m1]
            u = A.T @ r
            if mutate != "no_image_weight":
                u[m2] /= W_img[m2]
            x = x + gamma * u
            if return_all:
                out.append(x.copy())
    return np.array(out) if return_all else x


def sirt_dd_convention(A, b, iters, gamma=1.0, W_sino=None, W_img=None, mutate=None):
    """sys_mat/sirt_dd.py:46-72: x0 = 1e-5 everywhere, iterations
    i = 1 .. iters-1."""
    x0 = np.full(A.shape[1], np.float64(np.float32(1e-5)))
    return sirt(A, b, max(iters - 1, 0), gamma, x0, W_sino, W_img, mutate=mutate)


def sirt_test_convention(A, b, iters, gamma=1.0, W_sino=None, W_img=None, mutate=None):
    """tests/Test_SIRT/sirt_test.py:21-40 with beta = 0: x0 = 0, `iters`
    updates, all iterates (iters, n)."""
    return sirt(A, b, iters, gamma, None, W_sino, W_img, return_all=True, mutate=mutate)
