"""ORACLE: the Chambolle-Pock TV reconstruction of tv.py:229-294 in float64
with an explicit matrix A -- test infrastructure only.

  sigma_z = 1/grad_abs(1) = 1/2, sigma_q = 1/(A 1) (0 where A 1 = 0),
  tau = 1/(A^T 1 + div_abs(1)) = 1/(A^T 1 + 2*ndim)            (tv.py:257-259)
  z += sigma_z grad(ubar); z = z / max(1, |z|)                  (tv.py:278-279)
  q = (q + sigma_q (A ubar - g)) / (1 + sigma_q lam)            (tv.py:281-282)
  u' = u + tau (div z - A^T q); ubar = u' + theta (u' - u)       (tv.py:285-291)
  grad / div: periodic forward / backward differences (np.roll, tv.py:207-224).

pyT2 (tv.py:24-95) is meant to project each pixel's gradient vector onto the
unit ball (its singular-value clipping of a 1x2 matrix); as written it calls
np.min(array, array) and raises, so the projection is restated from its
intent.  The saddle point solves  min_u TV(u) + |A u - g|^2 / (2 lam),  TV
the isotropic periodic total variation.  Pinned in tests/test_pdhg_oracle.py:
div = -grad^T; with the TV term off the iteration reaches the least-squares
solution (its fixed point); with it on, the duality gap of that problem
closes (primal value = dual value, A^T q = div z, |z| <= 1).
"""
import numpy as np


def grad(u):
    return np.stack([np.roll(u, -1, c) - u for c in range(u.ndim)])


def div(z):
    return sum(z[c] - np.roll(z[c], 1, c) for c in range(z.shape[0]))


def pdhg(A, g, shape, iters, lam=1.0, theta=1.0, tv=True, return_all=False, return_duals=False):
    shape = tuple(shape)
    nd = len(shape)
    A1 = A @ np.ones(A.shape[1])
    sig_q = np.where(A1 != 0, 1.0 / np.where(A1 != 0, A1, 1.0), 0.0)
    tau = (1.0 / (A.T @ np.ones(A.shape[0]) + 2.0 * nd)).reshape(shape)
    g = np.asarray(g, np.float64).reshape(-1)
    u = np.zeros(shape)
    ubar = u.copy()
    z = np.zeros((nd,) + shape)
    q = np.zeros(A.shape[0])
    out = []
    for _ in range(iters):
        if tv:
            z = z + 0.5 * grad(ubar)
            nrm = np.sqrt((z * z).sum(0))
            z = z / np.maximum(1.0, nrm)
        q = (q + sig_q * (A @ ubar.reshape(-1) - g)) / (1.0 + sig_q * lam)
        un = u + tau * (div(z) - (A.T @ q).reshape(shape))
        ubar = un + theta * (un - u)
        u = un
        if return_all:
            out.append(u.copy())
    if return_duals:
        return u, z, q
    return np.array(out) if return_all else u
