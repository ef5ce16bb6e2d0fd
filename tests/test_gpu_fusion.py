"""Register transforms that share HBM passes (cvdvSetTransformFusion).

A register served by the four-step FFT with equal digits (2^10 .. 2^16 points
whose tile does not hold them whole) runs each transform as two half passes and
fuses the pending half of one transform with the first half of the next, the
phases between them applied in the same pass.  Checked here through the C ABI:

* against the oracle (FP64, 'cuda' convention) on circuits that keep the strided
  register's transforms chained: pending bilinear q1*q2 terms across both
  oscillators, conditional (qubit-controlled) terms, parity and swaps while a
  half is pending, and readouts that complete a pending half of either kind;
* fused against unfused (the whole-transform path) on the headline gate mix;
* that fusion happens: m chained transforms of the strided register take m + 1
  FFT launches instead of 2m.
Bars as in test_gpu_parity.py: complex128 1e-11 relative amplitude error and
1e-12 fidelity loss; mixed 5e-5 and 1e-6.
"""

import ctypes

import numpy as np
import pytest

from golden_cases import apply_ops, build_sep
from oracle.cvdv_oracle import CvdvOracle
from oracle.separable_np import SeparableStateNP

pytestmark = pytest.mark.gpu

PRECS = ["complex128", "mixed"]


def _loss(a, b):
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    return 1.0 - abs(np.vdot(a, b)) ** 2


def _bars(prec):
    return (1e-11, 1e-12) if prec == "complex128" else (5e-5, 1e-6)


def _chain_ops(rng, regs, n):
    """Gates on the strided register (index 1) and its neighbours: mostly transforms
    and phases, with the operations that must complete a pending half mixed in."""
    q = regs[0]
    two = len(regs) > 2
    ops = []
    for _ in range(n):
        k = int(rng.integers(0, 16))
        th = float(rng.uniform(-3.0, 3.0))
        a = [float(rng.uniform(-0.6, 0.6)), float(rng.uniform(-0.6, 0.6))]
        qb = int(rng.integers(0, q))
        rr = float(rng.uniform(0.05, 0.25)) * (1 if rng.random() < 0.5 else -1)
        table = [["d", 1, a], ["s", 1, rr], ["r", 1, th], ["ftQ2P", 1], ["ftP2Q", 1],
                 ["cd", 1, 0, qb, a], ["cr", 1, 0, qb, th], ["cs", 1, 0, qb, rr],
                 ["sheer", 1, rr], ["p", 1], ["h", 0, qb], ["cp", 1, 0, qb],
                 ["q1q2", 1, 2, 0.2 * th] if two else ["rz", 0, qb, th],
                 ["bs", 1, 2, th if regs[1] == regs[2] else 0.3 * th] if two else ["d", 1, a],
                 ["d", 2, a] if two else ["ftQ2P", 1],
                 ["swap", 1, 2] if two and regs[1] == regs[2] else ["ftP2Q", 1]]
        ops.append(table[k])
    return ops


def _run(regs, init, ops, prec, fuse=True, reads=()):
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv(regs, precision=prec)
    h._c("cvdvSetTransformFusion", ctypes.c_int(1 if fuse else 0))
    h.initStateVector(build_sep(SeparableState, regs, init, device="cuda"))
    apply_ops(h, ops)
    return h


# strided registers served by the fused halves: [1,10,4] (2^10 at stride 16: row
# blocks, per-butterfly twiddles), [2,12,6] (2^12 at stride 64: one row block),
# [1,10,10] (the headline's shape at 2^10), [2,14] complex128 (contiguous 2^14 in
# two passes: twiddles per column), [1,4,14] (complex128: the contiguous 2^14
# register's one-column tiles over several row blocks, with beam-splitter and
# q1q2 terms from the small oscillator)
SHAPES = [([1, 10, 4], 31), ([2, 12, 6], 32), ([1, 10, 10], 33), ([2, 14], 34), ([1, 4, 14], 35)]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("regs,seed", SHAPES, ids=["x".join(map(str, s[0])) for s in SHAPES])
def test_fused_chains_against_oracle(regs, seed, prec):
    print(f"[fusion] oracle parity {regs} seed {seed} {prec}", flush=True)
    rng = np.random.default_rng(seed)
    init = [["setUniform", 0]] + [["setCoherent", r, [0.3 * r, -0.4]] for r in range(1, len(regs))]
    ops = _chain_ops(rng, regs, 50)
    h = _run(regs, init, ops, prec)
    o = CvdvOracle(regs, convention="cuda")
    o.initStateVector(build_sep(SeparableStateNP, regs, init))
    apply_ops(o, ops)
    hs, os_ = h.getState(), o.getState()
    err = np.max(np.abs(hs - os_)) / np.max(np.abs(os_))
    fl = _loss(hs, os_)
    print(f"    rel max amp err {err:.3e}, fidelity loss {fl:.3e}", flush=True)
    ea, ef = _bars(prec)
    assert err <= ea and fl <= ef


@pytest.mark.parametrize("prec", PRECS)
def test_readouts_complete_pending_halves(prec):
    """Readouts after an odd (L half pending) and an even (twiddle + H half
    pending) number of chained transforms, and gates on the other registers whose
    phases read the pending register (bilinear terms) in between."""
    print(f"[fusion] readouts with pending halves {prec}", flush=True)
    regs = [1, 10, 10]
    init = [["setUniform", 0], ["setCoherent", 1, [0.5, 0.2]], ["setFock", 2, 1]]
    o = CvdvOracle(regs, convention="cuda")
    o.initStateVector(build_sep(SeparableStateNP, regs, init))
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv(regs, precision=prec)
    h.initStateVector(build_sep(SeparableState, regs, init, device="cuda"))
    ea, _ = _bars(prec)
    rtol = 1e-11 if prec == "complex128" else 2e-6
    steps = [
        [["ftQ2P", 1]],                                   # L half pending
        [["sheer", 1, 0.1], ["ftP2Q", 1]],                # fused; twiddle + H pending
        [["q1q2", 1, 2, 0.15], ["ftQ2P", 2]],             # other register reads the pending one
        [["ftQ2P", 1], ["cd", 1, 0, 0, [0.3, -0.2]]],     # conditional terms across the chain
        [["ftP2Q", 1], ["p", 1], ["ftQ2P", 1]],           # parity while a half is pending
        [["swap", 1, 2], ["ftP2Q", 2], ["ftQ2P", 1]],     # the pending field changes register
        [["h", 0, 0], ["ftQ2P", 1]],
    ]
    for i, ops in enumerate(steps):
        apply_ops(h, ops)
        apply_ops(o, ops)
        if i % 2 == 0:
            for regs_m in ([1], [2, 0], [0]):
                np.testing.assert_allclose(h.m(*regs_m), o.m(*regs_m), rtol=0, atol=rtol, err_msg=f"step {i}")
            assert abs(h.m() - o.m()) < rtol * 10
        else:
            hs, os_ = h.getState(), o.getState()
            assert np.max(np.abs(hs - os_)) / np.max(np.abs(os_)) <= ea, f"step {i}"
    hs, os_ = h.getState(), o.getState()
    assert np.max(np.abs(hs - os_)) / np.max(np.abs(os_)) <= ea


@pytest.mark.parametrize("prec", PRECS)
def test_fused_matches_unfused_on_gaussian_mix(prec):
    """bench.make_circuit()'s 200 gates at [12,12]: fused against whole transforms."""
    import bench
    import torch
    from hipcvdv import HipCvdv, SeparableState

    print(f"[fusion] fused vs unfused, gaussian mix [12,12] {prec}", flush=True)
    regs = [12, 12]
    states = []
    for fuse in (1, 0):
        h = HipCvdv(regs, precision=prec)
        h._c("cvdvSetTransformFusion", ctypes.c_int(fuse))
        sep = SeparableState(regs)
        sep.setCoherent(0, 1.2 - 0.4j)
        sep.setCoeffs(1, torch.from_numpy(bench.squeezed_vacuum(regs[1], 0.5)))
        h.initStateVector(sep)
        for op in bench.make_circuit():
            getattr(h, op[0])(*op[1:])
        states.append(h.getState())
    err = np.max(np.abs(states[0] - states[1])) / np.max(np.abs(states[1]))
    fl = _loss(states[0], states[1])
    print(f"    rel max amp err {err:.3e}, fidelity loss {fl:.3e}", flush=True)
    ea, ef = _bars(prec)
    assert err <= ea and fl <= ef


@pytest.mark.parametrize("prec", PRECS)
def test_chain_launch_counts(prec):
    """m alternating transforms of the strided register of [12,12] with phases
    between them: m + 1 FFT launches fused (the last completes at the readout),
    2m whole; the states agree."""
    from hipcvdv import HipCvdv, SeparableState

    print(f"[fusion] launch counts {prec}", flush=True)
    regs = [12, 12]
    m = 6
    out = {}
    for fuse in (1, 0):
        h = HipCvdv(regs, precision=prec)
        h._c("cvdvSetTransformFusion", ctypes.c_int(fuse))
        sep = SeparableState(regs)
        sep.setCoherent(0, 0.8 + 0.3j)
        sep.setCoherent(1, -0.5j)
        h.initStateVector(sep)
        h.synchronize()
        h._lib.cvdvProfileEnable(h.ctx, 1)
        for i in range(m):
            (h.ftQ2P if i % 2 == 0 else h.ftP2Q)(0)
            h.sheer(0, 0.05 * (i + 1))
        h.flush()
        h.synchronize()
        tot, mx = ctypes.c_double(0), ctypes.c_double(0)
        n = h._lib.cvdvProfileRead(h.ctx, 0, ctypes.byref(tot), ctypes.byref(mx))
        h._lib.cvdvProfileEnable(h.ctx, 0)
        out[fuse] = (n, h.getState())
    assert out[1][0] == m + 1, out[1][0]
    assert out[0][0] == 2 * m, out[0][0]
    err = np.max(np.abs(out[1][1] - out[0][1])) / np.max(np.abs(out[0][1]))
    assert err <= _bars(prec)[0]
