"""The FFT autotuner (cvdvFftAutotune) through the C ABI.

* It changes tile choices only, so a circuit of transforms and phases gives the
  same state after tuning as before (complex128 to 1e-12 of the largest
  amplitude, complex64 to 2e-6: transform fusion on or off rounds
  differently), and both match the oracle.
* It leaves the caller's state alone (the timing runs on a private context).
* A second context of the same shape reuses the timings: no plan is timed
  again, and the choices are the same.
"""

import numpy as np
import pytest

from golden_cases import apply_ops
from oracle.cvdv_oracle import CvdvOracle

pytestmark = pytest.mark.gpu

OPS = [["ftQ2P", 0], ["sheer", 0, 0.02], ["ftP2Q", 0], ["ftQ2P", 1], ["phaseCubic", 1, 0.003],
       ["ftP2Q", 1], ["q1q2", 0, 1, 0.05], ["ftQ2P", 0], ["ftP2Q", 0], ["d", 1, [0.3, -0.2]],
       ["bs", 0, 1, 0.7], ["s", 0, 0.2], ["r", 1, 0.9]]


def _bars(prec):
    return 1e-12 if prec == "complex128" else 2e-6


def _psi(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
@pytest.mark.parametrize("regs", [[10, 10], [12, 8]], ids=["10x10", "12x8"])
def test_tuned_plans_give_the_same_state(regs, prec):
    from hipcvdv import HipCvdv

    print(f"[autotune] {regs} {prec}", flush=True)
    psi = _psi(1 << sum(regs), sum(regs))
    plain = HipCvdv(regs, precision=prec)
    plain.initFromArray(psi)
    apply_ops(plain, OPS)
    a = plain.getState()

    tuned = HipCvdv(regs, precision=prec)
    tuned.initFromArray(psi)
    info = tuned.autotune()
    print(f"    {info}", flush=True)
    assert info["plans_timed"] >= 0
    assert all(r in info for r in range(len(regs)))
    # the caller's state is untouched by the timing
    np.testing.assert_allclose(tuned.getState(), plain_init(regs, prec, psi), rtol=0, atol=0)
    apply_ops(tuned, OPS)
    b = tuned.getState()
    assert np.max(np.abs(a - b)) <= _bars(prec) * np.max(np.abs(a))

    o = CvdvOracle(regs, convention="cuda")
    o.initFromArray(psi)
    apply_ops(o, OPS)
    exp = o.getState()
    tol = 1e-11 if prec == "complex128" else 2e-5
    assert np.max(np.abs(b - exp)) <= tol * np.max(np.abs(exp))

    again = HipCvdv(regs, precision=prec)
    again.initFromArray(psi)
    info2 = again.autotune()
    assert info2["plans_timed"] == 0
    for r in range(len(regs)):
        assert {k: info2[r][k] for k in ("row_elems", "tile_cap", "fuse")} == \
               {k: info[r][k] for k in ("row_elems", "tile_cap", "fuse")}


def plain_init(regs, prec, psi):
    from hipcvdv import HipCvdv

    h = HipCvdv(regs, precision=prec)
    h.initFromArray(psi)
    return h.getState()
