"""Full-size checks (2^28 amplitudes: the [14,14] and [1]*14 + [14] shapes of the
measured workloads) through properties that hold at any size, so no oracle run
of the full state is needed:

* the centred DFT maps a coherent state to a coherent state, FT|a> = |-i a>
  (FT = exp(-i pi n / 2), so FT|n> = (-i)^n |n>); on a grid of 2^14 points the
  aliasing and sampling errors are far below FP64 rounding, and the oracle
  shows the identity to 2e-16 at [9,10] (it fails for +i a);
* FT followed by its inverse is the identity, and every pass keeps the norm;
* the complex64 (mixed) and complex128 runs of the same circuit agree;
* marginals are consistent: each qubit's marginal is the joint marginal summed
  over the other qubits, and the joint marginal sums to the norm.

The strided register of [14,14] runs the two-pass FFT, the contiguous one the
one-row persistent tile: both production paths at their measured shape.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REGS = [14, 14]
A0, A1 = 1.2 - 0.4j, -0.7 + 0.9j


def _coherent(regs, alphas, prec=None, device="cuda"):
    from hipcvdv import HipCvdv, SeparableState

    sep = SeparableState(regs, device=device)
    for r, a in enumerate(alphas):
        sep.setCoherent(r, a)
    if prec is None:
        return sep
    h = HipCvdv(regs, precision=prec)
    h.initStateVector(sep)
    return h


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_full_size_ft_of_coherent_states(prec):
    print(f"[full] FT of coherent states at [14,14] {prec}", flush=True)
    h = _coherent(REGS, [A0, A1], prec)
    h.ftQ2P(0)  # strided: two-pass FFT
    h.ftQ2P(1)  # contiguous: one-row tile
    f = h.getFidelity(_coherent(REGS, [-1j * A0, -1j * A1]))
    print(f"    fidelity with |-i a0> x |-i a1>: 1 - {1 - f:.3e}", flush=True)
    assert abs(1.0 - f) <= (1e-12 if prec == "complex128" else 2e-6)
    # wrong sign of the transform: far from the +i coherent state
    assert h.getFidelity(_coherent(REGS, [1j * A0, 1j * A1])) < 0.5


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_full_size_ft_round_trip(prec):
    print(f"[full] FT round trip at [14,14] {prec}", flush=True)
    h = _coherent(REGS, [A0, A1], prec)
    n0 = h.m()
    for r in (0, 1):
        h.ftQ2P(r)
        h.ftP2Q(r)
    f = h.getFidelity(_coherent(REGS, [A0, A1]))
    print(f"    round-trip fidelity 1 - {1 - f:.3e}, norm drift {h.m() - n0:.3e}", flush=True)
    assert abs(1.0 - f) <= (1e-12 if prec == "complex128" else 2e-6)
    assert abs(h.m() - n0) <= (1e-12 if prec == "complex128" else 1e-6)


def _circuit(h):
    """One of each decomposition on both oscillators (every kind of FFT pass)."""
    h.d(0, 0.4 - 0.3j)
    h.r(1, 0.9)
    h.s(0, 0.2)
    h.bs(0, 1, 0.7)
    h.q1q2(0, 1, 0.05)
    h.phaseCubic(1, 0.002)
    h.r(0, 2.2)  # a quarter-turn branch
    h.d(1, -0.5 + 0.2j)


def test_full_size_mixed_matches_complex128():
    print("[full] mixed vs complex128 at [14,14]", flush=True)
    hm = _coherent(REGS, [A0, A1], "mixed")
    hd = _coherent(REGS, [A0, A1], "complex128")
    _circuit(hm)
    _circuit(hd)
    f = hm.fidelityWith(hd)
    print(f"    fidelity 1 - {1 - f:.3e}, norms {hm.m():.9f} {hd.m():.12f}", flush=True)
    assert abs(1.0 - f) <= 2e-6
    assert abs(hd.m() - 1.0) <= 1e-11
    assert abs(hm.m() - 1.0) <= 2e-5
    # the marginals of the first oscillator agree bin by bin
    np.testing.assert_allclose(hm.m(0), hd.m(0), rtol=0, atol=2e-6)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_full_size_state_transfer_marginals(prec):
    import bench
    from hipcvdv import HipCvdv, SeparableState

    print(f"[full] state-transfer marginals at [1]*14 + [14] {prec}", flush=True)
    regs, cat, ops = bench.make_c3()
    nq = len(regs) - 1
    h = HipCvdv(regs, precision=prec)
    sep = SeparableState(regs)
    for k in range(nq):
        sep.setZero(k)
    sep.setCat(nq, cat)
    h.initStateVector(sep)
    for op in ops[:12]:  # four rounds of cd / h / cr
        getattr(h, op[0])(*op[1:])
    norm = h.m()
    joint = np.asarray(h.m(*range(nq)))
    assert joint.shape == (2,) * nq
    tol = 1e-12 if prec == "complex128" else 1e-6
    assert abs(joint.sum() - norm) <= tol
    for k in range(nq):
        single = np.asarray(h.m(k))
        other = tuple(a for a in range(nq) if a != k)
        np.testing.assert_allclose(single, joint.sum(axis=other), rtol=0, atol=tol)
    # qubits 4..13 are untouched by the first four rounds: still |0>
    for k in range(4, nq):
        np.testing.assert_allclose(np.asarray(h.m(k)), [norm, 0.0], rtol=0, atol=tol)
