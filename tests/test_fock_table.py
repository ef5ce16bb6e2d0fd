"""The Fock-state table of the QFT precision sweep (bench._fock_table).

It runs the reference builder's Hermite recurrence (separable.py:62-74) on
mantissas with a per-point log scale, so that states above n ~ 745 keep their
tails (e^{-x^2/2} alone underflows for x^2/2 > 745).  Pins: equal to the
oracle's plain recurrence where that one does not underflow, and discretely
orthonormal far beyond that point on a 2^12-point grid (the grid-sampled
Hermite functions are orthonormal to rounding while they fit the grid, about
n < pi N / 4)."""

import math

import numpy as np

from bench import _fock_table
from oracle.separable_np import SeparableStateNP


def test_matches_plain_recurrence_below_underflow():
    N = 512
    F = _fock_table(N)
    sep = SeparableStateNP([9])
    for n in (0, 1, 2, 7, 50, 300):
        sep.setFock(0, n)
        np.testing.assert_allclose(F[n], sep.register_arrays[0].real, rtol=0, atol=1e-14)


def test_orthonormal_beyond_the_underflow_point():
    N = 4096
    F = _fock_table(N)
    m = int(0.9 * math.pi * N / 4)  # well above n = 745
    G = F[:m] @ F[:m].T
    assert np.max(np.abs(G - np.eye(m))) < 1e-11
    # the plain recurrence loses these states' tails: its norms drift from 1 above n ~ 745
    sep = SeparableStateNP([12])
    sep.setFock(0, 2000)
    plain = sep.register_arrays[0].real
    assert abs(np.sum(plain ** 2) - 1.0) > 1e-3
    assert abs(np.sum(F[2000] ** 2) - 1.0) < 1e-11
