"""The brick-gather backprojector's candidate index q -> (q / mV, q mod mV)
as floor((q + 0.5) r), r the hardware reciprocal of mV (vir_amd/csrc/siddon.hip,
k_siddon_bp_wave: ray_at), for q < 2^21.

The hardware reciprocal is within 1 ulp of 1/mV, so r is RN(1/mV) or one of
its float neighbours.  The argument: the exact (q + 0.5)/mV has its
fraction in [0.5/mV, 1 - 0.5/mV], and the two roundings move it by less than
(q + 0.5)/mV 2^-22 < 0.5/mV while q + 0.5 < 2^21.  Checked here in float32
for every divisor up to 1500 at the quotient boundaries (q = k mV - 1 and
k mV, where the fraction is closest to an integer), and at random q, for all
three candidate reciprocals; past 2^21 the same check must fail somewhere
(the bound is not loose by orders of magnitude)."""
import numpy as np

LIMIT = 1 << 21


def _quotients(q, m, r):
    q = np.asarray(q, np.float32)
    return np.floor((q + np.float32(0.5)) * r).astype(np.int64)


def _recips(m):
    r = np.float32(1.0) / np.float32(m)
    return [r, np.nextafter(r, np.float32(0.0)), np.nextafter(r, np.float32(np.inf))]


def test_quotient_at_boundaries_and_random():
    rng = np.random.default_rng(3)
    for m in range(1, 1501):
        k = np.arange(1, LIMIT // m + 1, dtype=np.int64)
        q = np.concatenate([k * m - 1, k * m, rng.integers(0, LIMIT, 200)])
        q = q[(q >= 0) & (q < LIMIT)]
        for r in _recips(m):
            assert np.array_equal(_quotients(q, m, r), q // m), (m, r)


def test_bound_is_needed():
    # beyond 2^21 some divisor and neighbour reciprocal gives a wrong quotient
    wrong = False
    for m in (3, 5, 7, 11, 13, 641, 1023):
        k = np.arange(LIMIT // m, (1 << 26) // m, dtype=np.int64)
        q = np.concatenate([k * m - 1, k * m])
        for r in _recips(m):
            wrong |= not np.array_equal(_quotients(q, m, r), q // m)
    assert wrong
