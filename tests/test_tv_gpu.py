"""GPU total-variation denoising (vir_tv_chambolle) against the float64
oracle (oracle/tv_ref.py, pinned by closed forms in test_tv_oracle.py), and
SIRT with the TV step (beta > 0, sirt_test.py:37-38) against the float64
recurrence with the oracle's denoiser after every update."""
import numpy as np
import pytest

from oracle import siddon as osd
from oracle import sirt_ref, tv_ref

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape,w", [((200,), 0.3), ((40, 33), 0.2), ((17, 12, 9), 0.15)])
def test_fixed_iterations_equal_the_oracle(shape, w):
    import torch
    from vir_amd.tv import denoise_tv_chambolle
    rng = np.random.default_rng(len(shape))
    g = rng.normal(size=shape).astype(np.float32)
    u = denoise_tv_chambolle(torch.from_numpy(g).cuda(), weight=w, eps=0, max_num_iter=100)
    assert tuple(u.shape) == shape and u.dtype == torch.float32
    assert rel(u.cpu().numpy(), tv_ref.chambolle(g, w, 100)) <= TOL
    # numpy in -> numpy out
    assert isinstance(denoise_tv_chambolle(g, weight=w, max_num_iter=5), np.ndarray)


@pytest.mark.parametrize("shape", [(24, 5), (24, 3, 4)])
def test_step_converges_to_the_closed_form(shape):
    from vir_amd.tv import denoise_tv_chambolle
    n1, n2, w = 10, 14, 0.9
    g = np.full(shape, 1.0, np.float32)
    g[:n1] = 0.2
    # every iteration (the dual change falls below any useful eps long before
    # the 3-D case converges, tau = 1/12)
    u = denoise_tv_chambolle(g, weight=w, eps=0, max_num_iter=20000)
    want = np.concatenate([np.full(n1, 0.2 + w / n1), np.full(n2, 1.0 - w / n2)])
    # float32 stalls a few 1e-5 from the minimiser (the fixed-iteration
    # parity against float64 above is the 1e-5 test)
    assert np.abs(u.reshape(24, -1) - want[:, None]).max() <= 5e-5


def test_stopping_rule_ends_early():
    import torch
    from vir_amd.tv import denoise_tv_chambolle
    g = torch.full((30, 30), 0.5, device="cuda")
    g[10:20, 10:20] = 1.0
    # a constant-gradient-free interior converges; eps large stops at the first check
    u10 = denoise_tv_chambolle(g, weight=0.1, eps=1e9, max_num_iter=200, check_every=10)
    ref = tv_ref.chambolle(g.cpu().numpy(), 0.1, 10)
    assert rel(u10.cpu().numpy(), ref) <= TOL


def test_sirt_with_tv_step_equals_the_float64_recurrence():
    from vir_amd import sirt as vsirt
    from vir_amd.proj_geom import geom_circular
    v = np.arange(30) * np.pi / 30
    d = (np.arange(41) - 20.0) * 1.0
    s, t = geom_circular(d, v, "par")
    nP = (28, 28, 1)
    ind, ln, off = osd.siddons(s, t, nP, 1.0, 0.0, flat=True)
    A = sirt_ref.csr_matrix(ind, ln, off, int(np.prod(nP)))
    rng = np.random.default_rng(6)
    x_true = np.zeros(nP)
    x_true[6:22, 8:20] = 1.0
    b = (A @ x_true.reshape(-1) + 0.05 * rng.normal(size=A.shape[0])).astype(np.float32)
    W_sino = (A @ np.ones(A.shape[1])).astype(np.float32)
    W_img = (A.T @ np.ones(A.shape[0])).astype(np.float32)
    beta, iters, tv_iters = 0.05, 6, 40
    sino = b.reshape(30, 41)
    got = vsirt.sirt(sino, (ind, ln, off), W_sino.reshape(30, 41), W_img.reshape(nP), nP, beta=beta,
                     iters=iters, tv_kw=dict(eps=0, max_num_iter=tv_iters))
    x = np.zeros(A.shape[1])
    for i in range(iters):
        x = sirt_ref.sirt(A, b, 1, x0=x, W_sino=W_sino, W_img=W_img)
        x = tv_ref.chambolle(x.reshape(nP), beta, tv_iters).reshape(-1)
        assert rel(got[..., i], x) <= TOL, i
