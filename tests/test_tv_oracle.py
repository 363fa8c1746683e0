"""Pins of oracle/tv_ref.py (Chambolle's TV denoising, float64) by the
mathematics of the ROF problem; each planted mistake fails one of them."""
import numpy as np
import pytest

from oracle import tv_ref


def test_div_is_minus_grad_transpose():
    rng = np.random.default_rng(0)
    for shape in ((17,), (9, 7), (5, 6, 4)):
        u = rng.normal(size=shape)
        q = rng.normal(size=(len(shape),) + shape)
        for a in range(len(shape)):         # grad's last index is zero: so is q's
            sl = [a] + [slice(None)] * len(shape)
            sl[1 + a] = -1
            q[tuple(sl)] = 0.0
        lhs = float((tv_ref.grad(u) * q).sum())
        rhs = -float((u * tv_ref.div(q)).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def _step(shape, n1, a=0.2, b=1.0):
    g = np.full(shape, b)
    g[:n1] = a
    return g


def _step_ok(u, n1, n2, a, b, w, tol):
    want = np.concatenate([np.full(n1, a + w / n1), np.full(n2, b - w / n2)])
    prof = u.reshape(u.shape[0], -1)
    return np.abs(prof - want[:, None]).max() <= tol


@pytest.mark.parametrize("shape", [(24,), (24, 5), (24, 3, 4)])
def test_step_denoises_to_the_closed_form_plateaus(shape):
    n1, n2, w = 10, 14, 0.9
    g = _step(shape, n1)
    u = tv_ref.chambolle(g, w, 16000)      # tau = 1/(4 ndim): slower in 2-D and 3-D
    assert _step_ok(u, n1, n2, 0.2, 1.0, w, 1e-9)


def test_large_weight_gives_the_mean():
    n1, n2 = 10, 14
    g = _step((24, 3), n1)
    w = 8.0          # well above (b - a) n1 n2 / (n1 + n2) = 4.67 (convergence is slow near it)
    u = tv_ref.chambolle(g, w, 16000)
    assert np.abs(u - g.mean()).max() <= 1e-9


def test_constant_image_is_a_fixed_point():
    g = np.full((6, 5, 4), 0.37)
    assert np.array_equal(tv_ref.chambolle(g, 0.5, 50), g)


@pytest.mark.parametrize("mutate", tv_ref.MUTATIONS)
def test_planted_mistakes_fail_the_pins(mutate):
    n1, n2, w = 10, 14, 0.9
    g = _step((24, 5), n1)
    u = tv_ref.chambolle(g, w, 16000, mutate=mutate)
    assert not (np.all(np.isfinite(u)) and _step_ok(u, n1, n2, 0.2, 1.0, w, 1e-4))
