"""ORACLE: total-variation denoising by Chambolle's projection algorithm,
float64 -- test infrastructure only.

The reference denoises SIRT iterates with scikit-image's
denoise_tv_chambolle (tests/Test_SIRT/sirt_test.py:13,38), which is not
installed here (parity against it unpinned).  This restates the published
algorithm: A. Chambolle, "An algorithm for total variation minimization and
applications", J. Math. Imaging Vis. 20 (2004) 89-97, eq. (9), for

    u = argmin_u  ||u - g||^2 / 2 + w TV(u),   TV(u) = sum_i |(grad u)_i|

with grad the forward difference along each axis (zero at the last index)
and div = -grad^T.  In the dual variable q = w p (|p| <= 1):

    q <- (q + tau grad(div q - g)) / (1 + (tau / w) |grad(div q - g)|)
    u  = g - div q

Pinned in tests/test_tv_oracle.py: div = -grad^T; a constant image is a
fixed point; a step of n1 samples at a and n2 at b (constant along the other
axes) denoises to the plateaus a + w/n1 and b - w/n2 (the exact discrete
minimiser while the jump stays positive); with w at or above
(b - a) n1 n2 / (n1 + n2) the image becomes its mean; planted mistakes fail.
"""
import numpy as np

MUTATIONS = ("div_sign", "no_norm", "grad_backward")


def grad(u):
    """Forward differences along every axis, zero at the last index:
    (ndim,) + u.shape."""
    u = np.asarray(u, np.float64)
    out = np.zeros((u.ndim,) + u.shape)
    for a in range(u.ndim):
        sl = [slice(None)] * u.ndim
        sl[a] = slice(0, -1)
        out[a][tuple(sl)] = np.diff(u, axis=a)
    return out


def div(q):
    """-grad^T: sum over axes of q_a(j) - q_a(j - e_a), with q_a(-1) = 0
    and q_a taken as 0 at the last index."""
    q = np.asarray(q, np.float64)
    nd = q.shape[0]
    out = np.zeros(q.shape[1:])
    for a in range(nd):
        qa = q[a].copy()
        sl = [slice(None)] * nd
        sl[a] = slice(-1, None)
        qa[tuple(sl)] = 0.0
        out += qa
        lo = [slice(None)] * nd
        hi = [slice(None)] * nd
        lo[a] = slice(1, None)
        hi[a] = slice(0, -1)
        out[tuple(lo)] -= qa[tuple(hi)]
    return out


def chambolle(g, weight, iters, tau=None, q0=None, return_q=False, mutate=None):
    """`iters` iterations of eq. (9) from q0 (default 0); returns u
    (and q)."""
    if mutate is not None and mutate not in MUTATIONS:
        raise ValueError(mutate)
    g = np.asarray(g, np.float64)
    nd = g.ndim
    tau = 1.0 / (4.0 * nd) if tau is None else tau
    q = np.zeros((nd,) + g.shape) if q0 is None else np.array(q0, np.float64)
    for _ in range(int(iters)):
        v = div(q) - g
        if mutate == "div_sign":
            v = -div(q) - g
        gr = grad(v)
        if mutate == "grad_backward":
            gr = -np.roll(gr, 1, axis=tuple(range(1, nd + 1)))
        nrm = np.sqrt((gr * gr).sum(0))
        if mutate == "no_norm":
            nrm = 0.0 * nrm
        q = (q + tau * gr) / (1.0 + (tau / weight) * nrm)
    u = g - div(q)
    return (u, q) if return_q else u
