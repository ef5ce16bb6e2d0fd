"""GPU parity: HipCvdv (through the C ABI of libcvdv_hip.so) against the oracle.

The oracle runs with convention='cuda' (no init renormalisation, m() = sum |psi|^2,
rint/round turn counts) - the behaviour of the reference's CudaCvdv that HipCvdv
replaces.  Bars:
  complex128: max |amplitude error| <= 1e-12 * max|psi|, readouts <= 1e-11;
  mixed (complex64 amplitudes, FP64 angles/twiddles/sums): fidelity loss <= 1e-6
  and max amplitude error <= 2e-5 * max|psi|.
"""

import numpy as np
import pytest

from golden_cases import apply_ops, build_sep, c, load
from oracle.cvdv_oracle import CvdvOracle
from oracle.separable_np import SeparableStateNP

pytestmark = pytest.mark.gpu

Z, CASES = load()
PRECS = ["complex128", "mixed"]


def _hip():
    from hipcvdv import HipCvdv, SeparableState
    return HipCvdv, SeparableState


def fidelity_loss(a, b):
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    return 1.0 - abs(np.vdot(a, b)) ** 2


def amp_tol(prec):
    return 1e-12 if prec == "complex128" else 2e-5


def read_tol(prec):
    return 1e-11 if prec == "complex128" else 2e-6


def run_pair(regs, init, ops, prec):
    HipCvdv, SeparableState = _hip()
    h = HipCvdv(regs, precision=prec)
    h.initStateVector(build_sep(SeparableState, regs, init, device="cuda"))
    apply_ops(h, ops)
    o = CvdvOracle(regs, convention="cuda")
    o.initStateVector(build_sep(SeparableStateNP, regs, init))
    apply_ops(o, ops)
    return h, o


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("case", CASES, ids=[cs["name"] for cs in CASES])
def test_golden_case_parity(case, prec):
    print(f"[gpu] golden {case['name']} {prec}", flush=True)
    h, o = run_pair(case["regs"], case["init"], case["ops"], prec)
    hs, os_ = h.getState(), o.getState()
    scale = np.max(np.abs(os_))
    err = np.max(np.abs(hs - os_))
    fl = fidelity_loss(hs, os_)
    print(f"    max amp err {err:.3e} (scale {scale:.3e}), fidelity loss {fl:.3e}")
    assert err <= amp_tol(prec) * max(scale, 1e-300)
    assert fl <= (1e-13 if prec == "complex128" else 1e-6)
    for rd in case["reads"]:
        meth, *args = rd
        if meth == "getFidelity":
            got = h.getFidelity(build_sep(_hip()[1], case["regs"], args[0], device="cuda"))
            exp = o.getFidelity(build_sep(SeparableStateNP, case["regs"], args[0]))
        else:
            got = getattr(h, meth)(*args)
            exp = getattr(o, meth)(*args)
        got, exp = np.asarray(got, dtype=np.float64), np.asarray(exp, dtype=np.float64)
        assert got.shape == exp.shape, rd
        tol = read_tol(prec) * max(1.0, float(np.max(np.abs(exp))))
        if meth == "getPhotonNumber" and prec == "mixed":
            tol = 1e-4 * max(1.0, abs(float(exp)))
        np.testing.assert_allclose(got, exp, rtol=0, atol=tol, err_msg=f"{case['name']} {rd}")


def _random_ops(rng, regs, n):
    """Mixed circuit over every gate kind; register 0 holds the control qubits."""
    ops = []
    cv = list(range(1, len(regs)))
    for _ in range(n):
        k = int(rng.integers(0, 20))
        t = int(rng.choice(cv))
        q = int(rng.integers(0, regs[0]))
        a = [float(rng.uniform(-0.8, 0.8)), float(rng.uniform(-0.8, 0.8))]
        th = float(rng.uniform(-3.5, 3.5))
        rr = float(rng.uniform(0.05, 0.3)) * (1 if rng.random() < 0.5 else -1)
        two = len(cv) > 1
        u, v = (cv[0], cv[1]) if two else (t, t)
        table = [["d", t, a], ["cd", t, 0, q, a], ["r", t, th], ["cr", t, 0, q, th], ["s", t, rr],
                 ["cs", t, 0, q, rr], ["h", 0, q], ["rx", 0, q, th], ["ry", 0, q, th], ["rz", 0, q, th],
                 ["p", t], ["cp", t, 0, q], ["sheer", t, rr], ["phaseCubic", t, 0.02 * rr],
                 ["ftQ2P", t], ["ftP2Q", t],
                 ["bs", u, v, th if regs[u] == regs[v] else 0.4 * th] if two else ["x", 0, q],
                 ["cbs", u, v, 0, q, th if regs[u] == regs[v] else 0.4 * th] if two else ["y", 0, q],
                 ["q1q2", u, v, 0.3 * th] if two else ["z", 0, q],
                 ["swap", u, v] if two and regs[u] == regs[v] else ["h", 0, q]]
        ops.append(table[k])
    return ops


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("regs,seed", [([2, 6, 6], 11), ([1, 9], 12), ([3, 5, 5], 13), ([2, 8, 4], 14),
                                       ([2, 11], 15), ([1, 7, 7], 16),
                                       # one-row FFT tiles of the last register (2^11; 2^13 persistent)
                                       # with bilinear terms to the other oscillator pending
                                       ([1, 4, 11], 17), ([1, 2, 13], 18),
                                       # a contiguous register above one tile: two-pass FFT whose
                                       # second pass has one column and carries the twiddle
                                       ([1, 15], 19),
                                       # 32 amplitudes: FFT tiles of 4 threads, fewer than the
                                       # phase program's entries (staged by one thread)
                                       ([1, 2, 2], 20),
                                       # two-pass FFTs of a 2^16 contiguous register (narrowed
                                       # pass-1 tiles) and of a strided 2^13 register
                                       ([1, 16], 21), ([1, 13, 2], 22),
                                       # a contiguous 2^14 register: one-row tiles in complex64,
                                       # the two-pass split in complex128 (256 KiB rows)
                                       ([1, 14], 23), ([3, 14], 24)])
def test_random_circuit_parity(regs, seed, prec):
    print(f"[gpu] random {regs} seed {seed} {prec}", flush=True)
    rng = np.random.default_rng(seed)
    init = [["setUniform", 0]] + [["setCoherent", r, [0.4 * r, -0.3]] for r in range(1, len(regs))]
    ops = _random_ops(rng, regs, 60)
    h, o = run_pair(regs, init, ops, prec)
    hs, os_ = h.getState(), o.getState()
    err = np.max(np.abs(hs - os_)) / np.max(np.abs(os_))
    fl = fidelity_loss(hs, os_)
    print(f"    rel max amp err {err:.3e}, fidelity loss {fl:.3e}")
    # complex128: the phase angles grow like x^3 (phaseCubic) with |x| up to
    # sqrt(pi N / 2), ~1e5 rad at N = 2^16, so two FP64 evaluations of the same
    # angle differ by ~1e-11 rad there; the bar scales with the largest register
    c128_bar = 1e-11 if max(regs) <= 13 else 1e-10
    assert err <= (c128_bar if prec == "complex128" else 5e-5)
    assert fl <= (1e-12 if prec == "complex128" else 1e-6)
    # complex64 amplitudes: probabilities carry the FP32 norm drift of ~200 passes
    mtol = read_tol(prec) if prec == "complex128" else 2e-5
    for sel in ([0], [len(regs) - 1], list(range(len(regs)))[::-1]):
        np.testing.assert_allclose(h.m(*sel), o.m(*sel), rtol=0, atol=mtol)
    # FP32 amplitudes accumulate a norm drift of a few 1e-8 per pass (~120 passes here)
    assert abs(h.m() - o.m()) <= (1e-11 if prec == "complex128" else 2e-5)


def _spill_case(kind):
    """Pending phase programs with many tabulated terms: more conditional tables
    than one kernel holds, or terms that share tables.

    A conditional displacement with an imaginary amplitude is a phase e^{i k x}
    on the target, conditioned on the control qubit: one tabulated term per
    (target, control) pair.  'fft': eight such terms on one oscillator, consumed
    by its FFT.  'qubit': seven terms controlled by one qubit (after an FFT left
    a global phase pending), consumed by a Hadamard on that qubit.  'host': six
    conditional terms and the global phase (no unconditional table to carry it).
    'bits': Rz on six qubits of two registers (one-bit terms on different bits,
    read through shared tables) and kicks on the oscillator controlled by each
    of them, consumed by the oscillator's FFT and by the final flush.
    """
    if kind == "bits":
        regs = [3, 3, 7]
        init = [["setUniform", 0], ["setUniform", 1], ["setCoherent", 2, [0.6, 0.3]]]
        ops = [["rz", r, q, 0.3 + 0.4 * q + r] for r in (0, 1) for q in range(3)]
        ops += [["cd", 2, r, q, [0.0, 0.1 + 0.05 * q + 0.2 * r]] for r in (0, 1) for q in range(3)]
        ops += [["ftQ2P", 2], ["rz", 0, 1, 0.7]]
        return regs, init, ops
    if kind == "fft":
        regs = [1] * 8 + [6]
        init = [["setUniform", q] for q in range(8)] + [["setCoherent", 8, [0.5, -0.2]]]
        ops = [["cd", 8, q, 0, [0.0, 0.15 * (q + 1)]] for q in range(8)] + [["ftQ2P", 8]]
    else:
        n = 7 if kind == "qubit" else 6
        regs = [1] + [3] * n
        init = [["setUniform", 0]] + [["setCoherent", r, [0.3, 0.1 * r]] for r in range(1, n + 1)]
        ops = [["ftQ2P", 1]] + [["cd", r, 0, 0, [0.0, 0.2 * r]] for r in range(1, n + 1)] + [["h", 0, 0]]
    return regs, init, ops


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("kind", ["fft", "qubit", "host", "bits"])
def test_spilled_phase_tables(kind, prec):
    print(f"[gpu] spilled phase tables {kind} {prec}", flush=True)
    regs, init, ops = _spill_case(kind)
    h, o = run_pair(regs, init, ops, prec)
    hs, os_ = h.getState(), o.getState()
    err = np.max(np.abs(hs - os_)) / np.max(np.abs(os_))
    print(f"    rel max amp err {err:.3e}")
    assert err <= amp_tol(prec)
    assert fidelity_loss(hs, os_) <= (1e-13 if prec == "complex128" else 1e-6)


@pytest.mark.parametrize("prec", PRECS)
def test_readouts_after_lazy_layout(prec):
    """Wigner/Husimi/marginals/fidelity while registers are swapped and reflected."""
    print(f"[gpu] lazy layout readouts {prec}", flush=True)
    regs = [1, 6, 6]
    init = [["setUniform", 0], ["setCoherent", 1, [0.7, 0.2]], ["setFock", 2, 2]]
    ops = [["bs", 1, 2, 2.9], ["p", 1], ["cd", 2, 0, 0, [0.3, -0.2]], ["swap", 1, 2], ["rz", 0, 0, 0.3]]
    h, o = run_pair(regs, init, ops, prec)
    for r in (1, 2):
        np.testing.assert_allclose(h.getWigner(r), o.getWigner(r), rtol=0, atol=read_tol(prec))
        np.testing.assert_allclose(h.getHusimiQ(r), o.getHusimiQ(r), rtol=0, atol=read_tol(prec))
        np.testing.assert_allclose(h.m(r, 0), o.m(r, 0), rtol=0, atol=read_tol(prec))
    HipCvdv, SeparableState = _hip()
    ref = HipCvdv(regs, precision="complex128")
    ref.initFromArray(o.getState())
    assert abs(h.fidelityWith(ref) - 1.0) <= (1e-12 if prec == "complex128" else 1e-6)
    assert abs(h.getPhotonNumber(2) - o.getPhotonNumber(2)) <= (1e-9 if prec == "complex128" else 1e-4)


def test_deterministic_marginals():
    HipCvdv, SeparableState = _hip()
    regs = [3, 9, 4]
    sep = SeparableState(regs)
    sep.setUniform(0)
    sep.setCoherent(1, 0.5 + 0.2j)
    sep.setFock(2, 1)
    outs = []
    for _ in range(2):
        h = HipCvdv(regs, precision="mixed")
        h.initStateVector(sep)
        h.cd(1, 0, 1, 0.4 + 0.1j)
        h.bs(1, 2, 0.3) if regs[1] == regs[2] else h.r(1, 0.7)
        outs.append((h.m(1, 0), h.m(), h.m(2)))
    for a, b in zip(*outs):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_invalid_arguments_raise():
    HipCvdv, SeparableState = _hip()
    h = HipCvdv([2, 4])
    sep = SeparableState([2, 4])
    sep.setZero(0)
    sep.setZero(1)
    h.initStateVector(sep)
    with pytest.raises(ValueError):
        h.h(5, 0)
    with pytest.raises(ValueError):
        h.swap(0, 1)
    with pytest.raises(ValueError):
        h.q1q2(1, 1, 0.1)
    with pytest.raises(ValueError):
        h.cd(1, 0, 7, 0.5)
    assert abs(h.m() - 1.0) < 1e-15  # state untouched
