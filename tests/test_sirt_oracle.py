"""Pins of oracle/sirt_ref.py (the SIRT recurrence every SIRT, serving and
PDHG-free reconstruction test compares against).

* fixed point: on an inconsistent, overdetermined, full-column-rank A the
  iterates converge to the weighted least-squares solution
  argmin ||R^1/2 (A x - b)||, R = diag(1/W_sino), computed by lstsq;
* one step by hand on a 3x3 system with a zero row and a zero column (the
  W == 0 branches of sirt_dd.py:23-33 / sirt_test.py:26,33);
* the two driver conventions (start value, number of updates);
* each planted mistake is rejected by at least one of the above.
"""
import numpy as np
import pytest

from oracle import sirt_ref


def _system(seed=0, m=40, n=12):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0, 1, (m, n)) * (rng.uniform(0, 1, (m, n)) < 0.6)
    b = A @ rng.uniform(0, 1, n) + 0.05 * rng.standard_normal(m)   # inconsistent
    return A, b


def _wls(A, b):
    w = 1.0 / np.sqrt(A @ np.ones(A.shape[1]))
    return np.linalg.lstsq(A * w[:, None], b * w, rcond=None)[0]


def test_converges_to_the_weighted_least_squares_solution():
    A, b = _system()
    assert np.linalg.matrix_rank(A) == A.shape[1]
    x = sirt_ref.sirt(A, b, 20000)
    ref = _wls(A, b)
    assert np.linalg.norm(A @ ref - b) > 0.1          # the data really are inconsistent
    assert np.max(np.abs(x - ref)) < 1e-8 * np.max(np.abs(ref)) + 1e-10


def _hand_system():
    # row 1 and column 2 are zero: W_sino[1] = 0 and W_img[2] = 0
    A = np.array([[1.0, 2.0, 0.0], [0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    b = np.array([5.0, 7.0, 6.0])
    x0 = np.array([0.0, 0.0, 0.25])
    return A, b, x0


def test_one_step_by_hand():
    A, b, x0 = _hand_system()
    # W_sino = (3, 0, 7), W_img = (4, 6, 0); A x0 = 0
    # r = (5/3, A x0[1] = 0, 6/7); A^T r = (5/3 + 18/7, 10/3 + 24/7, 0)
    # u = (89/21 / 4, 142/21 / 6, 0 (W = 0: unscaled)) ; x1 = x0 + 0.5 u
    x1 = sirt_ref.sirt(A, b, 1, gamma=0.5, x0=x0)
    assert np.allclose(x1, [0.5 * 89 / 84, 0.5 * 71 / 63, 0.25], rtol=0, atol=1e-15)


def test_two_steps_by_hand():
    A, b, x0 = _hand_system()
    x1 = x0 + np.array([89 / 84, 71 / 63, 0.0])
    Ax = A @ x1
    r = np.array([(5 - Ax[0]) / 3, Ax[1], (6 - Ax[2]) / 7])
    u = A.T @ r / np.array([4.0, 6.0, 1.0])
    x2 = sirt_ref.sirt(A, b, 2, x0=x0)
    assert np.allclose(x2, x1 + u, rtol=0, atol=1e-14)


def test_driver_conventions():
    A, b = _system(1, 20, 6)
    # sirt_dd.sirt: x0 = 1e-5 (float32), iters - 1 updates
    x = sirt_ref.sirt_dd_convention(A, b, 5)
    x0 = np.full(6, float(np.float32(1e-5)))
    assert np.array_equal(x, sirt_ref.sirt(A, b, 4, x0=x0))
    assert np.array_equal(sirt_ref.sirt_dd_convention(A, b, 1), x0)
    # sirt_test.sirt: x0 = 0, iters updates, each iterate kept
    its = sirt_ref.sirt_test_convention(A, b, 3)
    assert its.shape == (3, 6)
    assert np.array_equal(its[0], sirt_ref.sirt(A, b, 1))
    assert np.array_equal(its[2], sirt_ref.sirt(A, b, 3))


def _pins_pass(mutate):
    """True when the pins above accept the oracle with `mutate` planted."""
    A, b, x0 = _hand_system()
    x1 = sirt_ref.sirt(A, b, 1, gamma=0.5, x0=x0, mutate=mutate)
    if not np.allclose(x1, [0.5 * 89 / 84, 0.5 * 71 / 63, 0.25], rtol=0, atol=1e-15):
        return False
    A, b = _system()
    x = sirt_ref.sirt(A, b, 20000, mutate=mutate)
    if not np.all(np.isfinite(x)) or np.max(np.abs(x - _wls(A, b))) > 1e-6:
        return False
    A, b = _system(1, 20, 6)
    its = sirt_ref.sirt_test_convention(A, b, 3, mutate=mutate)
    return its.shape == (3, 6) and np.allclose(its[0], sirt_ref.sirt(A, b, 1), rtol=0, atol=0)


def test_unmutated_oracle_passes_the_pins():
    assert _pins_pass(None)


@pytest.mark.parametrize("mutate", sirt_ref.MUTATIONS)
def test_planted_mistakes_are_rejected(mutate):
    assert not _pins_pass(mutate)
