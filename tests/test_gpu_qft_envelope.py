"""Precision envelope of the register QFT against the reference's own error
analysis (ERROR-ANALYSIS.md:13-16, analysis/qft_err.py, analysis/_common.py:67-86).

Every Fock state |n> of an N-point register goes through ftQ2P and is compared
with the analytic (-i)^n |n>; eps_n is the L2 error, swept up to the first n
with eps_n^2 >= 0.1.  The reference fits log eps <= a k + b per N over the
points with eps >= 1e-9 (least squares, b raised by the largest residual) and
publishes, for complex128, a(N) = 4.2125 N^-1/2 + 0.1027 and
b(N) = -5.7903 N^1/2 + 16.2724 (N = 32 .. 1024).

* complex128: each per-N fit lies on the published curves, within the scatter
  of the reference's own per-N points about them (|da| <= 0.01, |db| <= 4);
  the sweep stops where the grid runs out, near n = pi N / 4.
* mixed (complex64 amplitudes): the rounding floor (median eps over n < 8) and
  the largest distance from complex128 stay at ~1e-7.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _a_ref(N):
    return 4.2125 * N ** -0.5 + 0.1027


def _b_ref(N):
    return -5.7903 * N ** 0.5 + 16.2724


def _eps(q, prec):
    from bench import _fock_table
    from hipcvdv import HipCvdv

    N = 1 << q
    F = _fock_table(N)
    h = HipCvdv([q, q], precision=prec)
    h.initFromArray(F.astype(np.complex128).reshape(-1))
    h.ftQ2P(1)
    got = h.getState().reshape(N, N)
    target = F * ((-1j) ** (np.arange(N) % 4))[:, None]
    return np.sqrt(np.sum(np.abs(got - target) ** 2, axis=1)), got


@pytest.mark.parametrize("q", [6, 7, 8, 9])
def test_complex128_envelope_matches_reference_fit(q):
    from bench import _upper_fit

    N = 1 << q
    print(f"[qft envelope] complex128 N={N}", flush=True)
    eps, _ = _eps(q, "complex128")
    stop = int(np.argmax(eps ** 2 >= 0.1))
    assert eps[stop] ** 2 >= 0.1 and abs(stop - math.pi * N / 4) <= 0.05 * N
    ks = [k for k in range(stop + 1) if eps[k] >= 1e-9]
    a, b = _upper_fit(ks, eps[ks])
    print(f"    stop {stop}, a {a:.4f} (published {_a_ref(N):.4f}), b {b:.2f} (published {_b_ref(N):.2f})")
    assert abs(a - _a_ref(N)) <= 0.01
    assert abs(b - _b_ref(N)) <= 4.0
    assert np.median(eps[:8]) < 1e-14


@pytest.mark.parametrize("q", [6, 9])
def test_mixed_floor_and_distance_from_complex128(q):
    N = 1 << q
    print(f"[qft envelope] mixed N={N}", flush=True)
    e64, s64 = _eps(q, "complex128")
    e32, s32 = _eps(q, "mixed")
    stop = int(np.argmax(e64 ** 2 >= 0.1))
    d = np.sqrt(np.sum(np.abs(s32 - s64) ** 2, axis=1))[:stop]
    print(f"    mixed floor {np.median(e32[:8]):.2e}, max distance from complex128 {d.max():.2e}")
    assert np.median(e32[:8]) < 3e-7
    assert d.max() < 3e-7
