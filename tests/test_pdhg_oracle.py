"""Pins of the Chambolle-Pock oracle (oracle/pdhg_ref.py)."""
import numpy as np

from oracle import pdhg_ref


def test_div_is_minus_grad_transpose():
    rng = np.random.default_rng(0)
    for shape in ((7, 5), (4, 6, 3)):
        u = rng.normal(size=shape)
        z = rng.normal(size=(len(shape),) + shape)
        assert abs((pdhg_ref.grad(u) * z).sum() + (u * pdhg_ref.div(z)).sum()) < 1e-10


def test_without_tv_the_fixed_point_is_least_squares():
    rng = np.random.default_rng(1)
    shape = (4, 5)
    A = rng.uniform(0, 1, (40, 20))
    x = rng.uniform(0, 1, 20)
    g = A @ x + 0.01 * rng.normal(size=40)
    u = pdhg_ref.pdhg(A, g, shape, 4000, lam=1.0, tv=False)
    ls = np.linalg.lstsq(A, g, rcond=None)[0].reshape(shape)
    assert np.abs(u - ls).max() < 1e-6


def test_tv_projection_and_smoothing():
    rng = np.random.default_rng(2)
    shape = (12, 12)
    A = np.eye(144)
    clean = np.zeros(shape)
    clean[3:9, 4:10] = 1.0
    g = (clean + 0.3 * rng.normal(size=shape)).reshape(-1)
    u = pdhg_ref.pdhg(A, g, shape, 300, lam=0.5)
    tv = lambda a: np.abs(pdhg_ref.grad(a)).sum()
    assert tv(u) < 0.5 * tv(g.reshape(shape))
    assert np.linalg.norm(u - clean) < np.linalg.norm(g.reshape(shape) - clean)


def test_duality_gap_closes():
    """min_u TV(u) + |Au - g|^2/(2 lam) has the dual
    max { -lam/2 |q|^2 - <q, g> : A^T q = div z, |z| <= 1 }; at the saddle
    point both values agree.  A wrong sign, step or ball projection leaves a
    gap or violates the constraint."""
    rng = np.random.default_rng(3)
    shape, lam = (6, 7), 0.7
    A = rng.uniform(0, 1, (60, 42))
    clean = np.zeros(shape)
    clean[1:4, 2:6] = 1.0
    g = A @ clean.reshape(-1) + 0.2 * rng.normal(size=60)
    u, z, q = pdhg_ref.pdhg(A, g, shape, 20000, lam=lam, return_duals=True)
    tv = np.sqrt((pdhg_ref.grad(u) ** 2).sum(0)).sum()
    primal = tv + ((A @ u.reshape(-1) - g) ** 2).sum() / (2 * lam)
    dual = -0.5 * lam * (q * q).sum() - q @ g
    assert np.sqrt((z * z).sum(0)).max() <= 1 + 1e-12
    assert np.abs((A.T @ q).reshape(shape) - pdhg_ref.div(z)).max() < 1e-6
    assert abs(primal - dual) < 1e-6 * abs(primal)
