"""Native-grid phase-space readouts beyond the golden cases, through the C ABI:

* cvdvGetHusimiQWigner (the Wigner -> 2D FFT route, bosonic.cu:63-101) against the
  oracle's restatement (getHusimiQWigner, itself pinned by the coherent-state
  closed form in test_oracle_closed_form.py), traced over several slices;
* cvdvGetHusimiQ's route choice (bosonic.cu:13-25) forced both ways with
  cvdvSetFreeMemoryOverride: the FFT2 route when the overlap estimate exceeds
  80% of the (overridden) free memory, the overlap route otherwise;
* the device-output readouts equal the host-output ones bit for bit.
Bars: complex128 1e-11 of the largest value, mixed 2e-6 (amplitudes rounded to
complex64 once per pass; the readout itself runs in FP64).
"""

import ctypes

import numpy as np
import pytest

from golden_cases import apply_ops, build_sep
from oracle.cvdv_oracle import CvdvOracle
from oracle.separable_np import SeparableStateNP

pytestmark = pytest.mark.gpu

CASES = [
    ([8], [["setCoherent", 0, [1.0, 0.5]]], []),
    ([2, 8], [["setUniform", 0], ["setCoherent", 1, [-0.6, 0.4]]],
     [["cd", 1, 0, 0, [0.8, 0.0]], ["cr", 1, 0, 1, 0.7], ["s", 1, 0.2]]),
    ([7, 1], [["setFock", 0, 3], ["setUniform", 1]], [["cd", 0, 1, 0, [0.0, 0.9]], ["r", 0, 0.4]]),
]


def _pair(regs, init, ops, prec):
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv(regs, precision=prec)
    h.initStateVector(build_sep(SeparableState, regs, init, device="cuda"))
    apply_ops(h, ops)
    o = CvdvOracle(regs, convention="cuda")
    o.initStateVector(build_sep(SeparableStateNP, regs, init))
    apply_ops(o, ops)
    return h, o


def _tol(prec):
    return 1e-11 if prec == "complex128" else 2e-6


def _reg(regs):
    return max(range(len(regs)), key=lambda r: regs[r])


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_husimi_fft2_route_against_oracle(case, prec):
    regs, init, ops = CASES[case]
    print(f"[phase space] husimi fft2 route {regs} {prec}", flush=True)
    h, o = _pair(regs, init, ops, prec)
    r = _reg(regs)
    got, exp = h.getHusimiQWigner(r), o.getHusimiQWigner(r)
    assert np.max(np.abs(got - exp)) <= _tol(prec) * np.max(np.abs(exp))
    assert np.all(got >= 0)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_husimi_route_choice(prec):
    regs, init, ops = CASES[1]
    print(f"[phase space] husimi route choice {prec}", flush=True)
    h, o = _pair(regs, init, ops, prec)
    r = _reg(regs)
    # overlap estimate of [2,8]: 2^8*16 + 2*2^16*8 + 4*2^8*16 + 4*2^16*16 bytes ~ 5.3 MB
    h._c("cvdvSetFreeMemoryOverride", ctypes.c_size_t(1 << 20))
    tight = h.getHusimiQ(r)
    np.testing.assert_array_equal(tight, h.getHusimiQWigner(r))
    h._c("cvdvSetFreeMemoryOverride", ctypes.c_size_t(1 << 40))
    roomy = h.getHusimiQ(r)
    exp = o.getHusimiQ(r)
    assert np.max(np.abs(roomy - exp)) <= _tol(prec) * np.max(np.abs(exp))
    h._c("cvdvSetFreeMemoryOverride", ctypes.c_size_t(0))  # the device's own figure: overlap here
    np.testing.assert_array_equal(h.getHusimiQ(r), roomy)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_device_readouts_equal_host(prec):
    regs, init, ops = CASES[1]
    print(f"[phase space] device readouts {prec}", flush=True)
    h, _ = _pair(regs, init, ops, prec)
    r = _reg(regs)
    for name in ("getWigner", "getHusimiQ"):
        for bound in (None, 3.0):
            host = getattr(h, name)(r, bound)
            dev = getattr(h, name)(r, bound, device=True)
            assert dev.is_cuda and tuple(dev.shape) == host.shape
            np.testing.assert_array_equal(dev.cpu().numpy(), host)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
@pytest.mark.parametrize("regs,sel", [([2, 2, 12], [0, 1, 2]), ([2, 2, 12], [2]), ([1, 10], [0]),
                                      ([1, 1, 1, 1, 10], [0, 1, 2, 3]), ([3, 4, 5], [2, 0]),
                                      ([1] * 6 + [8], list(range(6))), ([1] * 6 + [8], [5, 3, 1, 0, 2, 4]),
                                      ([2, 3, 2, 6], [3, 1]), ([2, 2, 12], [2, 0, 1]), ([12], [0])],
                         ids=["2x2x12-012", "2x2x12-2", "1x10-0", "1x1x1x1x10-0123", "3x4x5-20",
                              "1x6+8-012345", "1x6+8-531024", "2x3x2x6-31", "2x2x12-201", "12-0"])
def test_library_reduction_comparator(regs, sel, prec):
    """cvdvMeasureMultipleCT (one hipTensor reduction of |psi|^2) against the own
    deterministic reduction, on a random state with a lazy layout pending."""
    from hipcvdv import HipCvdv, _lib

    print(f"[phase space] hipTensor marginal {regs} {sel} {prec}", flush=True)
    rng = np.random.default_rng(sum(regs))
    n = 1 << sum(regs)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    h = HipCvdv(regs, precision=prec)
    h.initFromArray(psi / np.linalg.norm(psi))
    h.p(len(regs) - 1)
    own = h.m(*sel)
    out = np.zeros(own.size)
    sel_c = (ctypes.c_int * len(sel))(*sel)
    h._lib.cvdvMeasureMultipleCT(h.ctx, sel_c, len(sel), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _lib.check()
    lib = out.reshape(own.shape[::-1]).T if own.ndim > 1 else out
    np.testing.assert_allclose(lib, own, rtol=1e-12, atol=1e-15)
