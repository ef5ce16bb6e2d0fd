"""cvdvMeasureMultipleNoAlloc (fixed device scratch, no allocation) against
cvdvMeasureMultiple, both called through the C ABI on the same states.

Both run the same deterministic reduction; where the fixed scratch holds the
split count cvdvMeasureMultiple chooses (every case below), the outputs are
bitwise equal.  Two calls in a row, the second on a smaller state, check that
the shared array keeps no state a call depends on; a marginal beyond the
array's 2^20 bins is refused with an argument error.
"""

import ctypes
from ctypes import POINTER, c_double, c_int

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _marginal(h, name, regs):
    from hipcvdv import _lib

    lib = _lib.get_lib()
    n = int(np.prod([h.register_dims[r] for r in regs]))
    out = np.full(n, np.nan)
    getattr(lib, name)(h.ctx, (c_int * len(regs))(*regs), c_int(len(regs)), out.ctypes.data_as(POINTER(c_double)))
    _lib.check()
    return out


def _random_state(shape, seed, prec):
    from hipcvdv import HipCvdv

    rng = np.random.default_rng(seed)
    n = 1 << sum(shape)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    h = HipCvdv(shape, precision=prec)
    h.initFromArray(psi)
    return h


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_no_alloc_matches_measure_multiple(prec):
    cases = [
        ([2, 3, 14], [[0, 1, 2], [2], [1, 0], [0]]),  # 2^19 amplitudes; 2^19 bins, 2^14 bins, 32, 4
        ([1, 10], [[1], [0], [0, 1]]),                 # smaller state and fewer bins on the second call
    ]
    for seed, (shape, sels) in enumerate(cases):
        print(f"[no-alloc] {prec} shape {shape}", flush=True)
        h = _random_state(shape, 100 + seed, prec)
        # a few gates so the state carries a pending program and a lazy layout
        h.h(0, 0)
        h.p(len(shape) - 1)
        h.d(len(shape) - 1, 0.3 - 0.2j)
        for regs in sels:
            a = _marginal(h, "cvdvMeasureMultiple", regs)
            b = _marginal(h, "cvdvMeasureMultipleNoAlloc", regs)
            assert np.array_equal(a, b), (shape, regs, np.max(np.abs(a - b)))
            assert abs(b.sum() - h.m()) < 1e-12


def test_no_alloc_refuses_oversize():
    print("[no-alloc] oversize marginal", flush=True)
    h = _random_state([21], 7, "mixed")
    with pytest.raises(ValueError):
        _marginal(h, "cvdvMeasureMultipleNoAlloc", [0])
    # the ordinary entry point still serves it, and the fixed one serves a fitting one after
    assert abs(_marginal(h, "cvdvMeasureMultiple", [0]).sum() - 1.0) < 1e-5
    h2 = _random_state([1, 12], 8, "complex128")
    a = _marginal(h2, "cvdvMeasureMultiple", [1, 0])
    b = _marginal(h2, "cvdvMeasureMultipleNoAlloc", [1, 0])
    assert np.array_equal(a, b)
