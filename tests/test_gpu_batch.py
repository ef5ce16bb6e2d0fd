"""Batched circuits (one launch, one workgroup per circuit) against the same
circuits run one by one through HipCvdv, and against the oracle.

The per-circuit path is itself pinned to the oracle (test_gpu_parity.py), so
agreement with it to FFT rounding (complex128: 1e-11 absolute; mixed: 2e-5)
plus direct oracle checks on a few complex128 circuits pin the batch engine.
Across launch shapes (threads, LDS budget, LDS-resident vs global work area)
the per-element arithmetic is identical, so those outputs must agree bit for bit.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GATES = ("h", "rx", "ry", "rz", "cd", "r", "s", "cp", "d", "cr", "p", "sheer")


def _gate_list(rng, shape, n):
    """Random gates for a circuit [q, nb] (qubits register 0, oscillator register 1),
    or [q, nb1, nb2] with a second oscillator (adds bs, q1q2, swap when equal)."""
    q = shape[0]
    cvs = list(range(1, len(shape)))
    out = []
    kinds = list(GATES)
    if len(cvs) == 2:
        kinds += ["bs", "q1q2"] + (["swap"] if shape[1] == shape[2] else [])
    for _ in range(n):
        k = kinds[int(rng.integers(0, len(kinds)))]
        qb = int(rng.integers(0, q))
        reg = cvs[int(rng.integers(0, len(cvs)))]
        th = float(rng.uniform(-math.pi, math.pi))
        a = complex(*(0.4 * rng.uniform(-1, 1, 2)))
        if k == "h":
            out.append(("h", 0, qb))
        elif k in ("rx", "ry", "rz"):
            out.append((k, 0, qb, th))
        elif k == "cd":
            out.append(("cd", reg, 0, qb, a))
        elif k == "d":
            out.append(("d", reg, a))
        elif k == "r":
            out.append(("r", reg, th))
        elif k == "cr":
            out.append(("cr", reg, 0, qb, th))
        elif k == "s":
            out.append(("s", reg, float(rng.uniform(-0.3, 0.3))))
        elif k == "cp":
            out.append(("cp", reg, 0, qb))
        elif k == "p":
            out.append(("p", reg))
        elif k == "sheer":
            out.append(("sheer", reg, float(rng.uniform(-0.2, 0.2))))
        elif k == "bs":
            t = th if shape[1] == shape[2] else 0.4 * th
            out.append(("bs", 1, 2, t))
        elif k == "q1q2":
            out.append(("q1q2", 1, 2, float(rng.uniform(-0.2, 0.2))))
        else:
            out.append(("swap", 1, 2))
    return out


def _init_arrays(rng, shape):
    import bench

    arrs = [np.zeros(1 << shape[0], dtype=np.complex128)]
    arrs[0][0] = 1.0
    for nb in shape[1:]:
        alpha = complex(*rng.uniform(-1.5, 1.5, 2))
        arrs.append(bench.coherent(nb, alpha).astype(np.complex128))
    return arrs


# shapes chosen to cover LDS-resident and global-work-area circuits, small and
# large FFT fields, three registers and a qubit-only-sized oscillator
SHAPES = [[1, 8], [2, 9], [3, 10], [1, 12], [3, 12], [2, 11], [1, 1], [2, 5, 5], [1, 6, 7], [3, 4, 4],
          [2, 12], [1, 3]]


def _make(prec, seed=7, n_gates=40, shapes=SHAPES):
    from hipcvdv import HipCvdvBatch

    rng = np.random.default_rng(seed)
    circuits = [(s, _init_arrays(rng, s), _gate_list(rng, s, n_gates)) for s in shapes]
    b = HipCvdvBatch([c[0] for c in circuits], precision=prec)
    for i, (_, _, gates) in enumerate(circuits):
        sim = b.circuit(i)
        for g in gates:
            getattr(sim, g[0])(*g[1:])
    b.initFromSeparable([c[1] for c in circuits])
    return b, circuits


def _sequential(prec, circuits):
    import torch
    from hipcvdv import HipCvdv, SeparableState

    outs = []
    for shape, arrs, gates in circuits:
        h = HipCvdv(shape, precision=prec)
        sep = SeparableState(shape)
        for r, a in enumerate(arrs):
            sep.setCoeffs(r, torch.from_numpy(a))
        h.initStateVector(sep)
        for g in gates:
            getattr(h, g[0])(*g[1:])
        outs.append((h.getState(), h.m(0), h.m()))
    return outs


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_batch_matches_sequential(prec):
    print(f"[batch] sequential parity {prec}", flush=True)
    b, circuits = _make(prec)
    b.setLaunch(256, 32768)
    b.run()
    probs = b.measure(0)
    norms = b.norms()
    seq = _sequential(prec, circuits)
    tol = 1e-11 if prec == "complex128" else 2e-5
    for i, (shape, _, _) in enumerate(circuits):
        st = b.getState(i)
        ref, m0, nrm = seq[i]
        assert st.shape == ref.shape
        err = np.max(np.abs(st - ref))
        assert err < tol, (i, shape, err)
        assert np.max(np.abs(probs[i] - m0)) < tol, (i, shape)
        assert abs(norms[i] - nrm) < tol, (i, shape)
        assert abs(norms[i] - np.sum(np.abs(st) ** 2)) < 1e-12


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_batch_tabulated_programs_match_sequential(prec):
    """The largest LDS budget: circuits of up to 2^15 amplitudes whose ops' phase
    programs (one polynomial group with up to three control qubits) are tabulated
    in the LDS left free beside the state or the staged FFT chunk; against the
    same circuits run one by one."""
    print(f"[batch] tabulated programs vs sequential {prec}", flush=True)
    shapes = [[3, 12], [2, 12], [3, 10], [1, 12], [3, 11], [2, 9]]
    rng = np.random.default_rng(21)
    circuits = []
    for s in shapes:
        gates = []
        for g in _gate_list(rng, s, 50):
            gates.append(g)
            if g[0] in ("d", "r", "s", "sheer"):  # controlled phases on several qubits
                qb = int(rng.integers(0, s[0]))
                gates.append(("cd", 1, 0, qb, complex(*(0.3 * rng.uniform(-1, 1, 2)))))
        circuits.append((s, _init_arrays(rng, s), gates))
    from hipcvdv import HipCvdvBatch

    b = HipCvdvBatch([c[0] for c in circuits], precision=prec)
    for i, (_, _, gates) in enumerate(circuits):
        sim = b.circuit(i)
        for g in gates:
            getattr(sim, g[0])(*g[1:])
    b.initFromSeparable([c[1] for c in circuits])
    b.setLaunch(512, 147456)
    b.run()
    seq = _sequential(prec, circuits)
    tol = 1e-11 if prec == "complex128" else 2e-5
    for i, (shape, _, _) in enumerate(circuits):
        err = np.max(np.abs(b.getState(i) - seq[i][0]))
        assert err < tol, (i, shape, err)


def test_batch_matches_oracle_c128():
    from oracle.cvdv_oracle import CvdvOracle
    from oracle.separable_np import SeparableStateNP

    print("[batch] oracle parity complex128", flush=True)
    shapes = [[1, 8], [2, 6, 6], [3, 9]]
    b, circuits = _make("complex128", seed=11, n_gates=30, shapes=shapes)
    b.run()
    for i, (shape, arrs, gates) in enumerate(circuits):
        o = CvdvOracle(shape, convention="cuda")
        sep = SeparableStateNP(shape)
        for r, a in enumerate(arrs):
            sep.setCoeffs(r, a)
        o.initStateVector(sep)
        for g in gates:
            getattr(o, g[0])(*g[1:])
        err = np.max(np.abs(b.getState(i) - o.getState()))
        assert err < 1e-11, (shape, err)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_batch_launch_shapes_bitwise(prec):
    print(f"[batch] launch shapes {prec}", flush=True)
    b, circuits = _make(prec, seed=3, n_gates=25)
    ref = None
    for threads, lds in [(256, 0), (256, 16384), (512, 32768), (256, 131072), (512, 65536), (512, 147456)]:
        print(f"[batch]   threads={threads} lds={lds}", flush=True)
        b.setLaunch(threads, lds)
        b.run()
        states = [b.getState(i) for i in range(len(circuits))]
        if ref is None:
            ref = states
        else:
            for i in range(len(circuits)):
                assert np.array_equal(states[i], ref[i]), (threads, lds, circuits[i][0])


def test_batch_rerun_and_extend():
    print("[batch] rerun and extend", flush=True)
    b, circuits = _make("complex128", seed=5, n_gates=10, shapes=[[1, 8], [2, 10]])
    b.run()
    s0 = [b.getState(i) for i in range(2)]
    b.run()
    for i in range(2):
        assert np.array_equal(b.getState(i), s0[i])
    # extend circuit 1 only; circuit 0 is unchanged, circuit 1 matches a fresh run
    b.circuit(1).r(1, 0.3)
    b.circuit(1).rz(0, 1, 0.2)
    b.run()
    assert np.array_equal(b.getState(0), s0[0])
    seq = _sequential("complex128", [(circuits[1][0], circuits[1][1], circuits[1][2] + [("r", 1, 0.3), ("rz", 0, 1, 0.2)])])
    assert np.max(np.abs(b.getState(1) - seq[0][0])) < 1e-11


def test_batch_extend_with_layout_only_gates():
    """Parity and swap change only the lazy layout of a recording: a rerun after
    extending a circuit with them (right after an FFT, so the pending program is
    unchanged) must recompile, not replay the previous launch."""
    print("[batch] extend with parity / swap", flush=True)
    shapes = [[1, 8], [2, 6, 6]]
    b, circuits = _make("complex128", seed=13, n_gates=8, shapes=shapes)
    b.run()
    extra = {0: [("ftQ2P", 1), ("p", 1)], 1: [("ftQ2P", 1), ("swap", 1, 2), ("ftP2Q", 2), ("p", 2)]}
    for i, ops in extra.items():
        for g in ops:
            getattr(b.circuit(i), g[0])(*g[1:])
    b.run()
    seq = _sequential("complex128", [(c[0], c[1], c[2] + extra[i]) for i, c in enumerate(circuits)])
    for i in range(len(shapes)):
        assert np.max(np.abs(b.getState(i) - seq[i][0])) < 1e-11, shapes[i]
    # one more layout-only gate: the pending program is still unchanged
    b.circuit(1).swap(1, 2)
    b.run()
    seq = _sequential("complex128", [(circuits[1][0], circuits[1][1], circuits[1][2] + extra[1] + [("swap", 1, 2)])])
    assert np.max(np.abs(b.getState(1) - seq[0][0])) < 1e-11


def test_batch_autotune_keeps_results():
    print("[batch] autotune", flush=True)
    b, circuits = _make("mixed", seed=9, n_gates=20)
    b.setLaunch(256, 0)
    b.run()
    before = [b.getState(i) for i in range(len(circuits))]
    us = b.autotune(reps=1)
    assert us > 0
    threads, lds = b.getLaunch()
    assert threads in (256, 512) and lds >= 0
    b.run()
    for i in range(len(circuits)):
        assert np.array_equal(b.getState(i), before[i])


def test_batch_errors():
    from hipcvdv import HipCvdvBatch

    print("[batch] errors", flush=True)
    with pytest.raises(ValueError):
        HipCvdvBatch([[1, 13]])  # register above 12 qubits
    with pytest.raises(ValueError):
        HipCvdvBatch([[10, 11]])  # 21 qubits
    b = HipCvdvBatch([[1, 8], [2, 9]], precision="mixed")
    with pytest.raises(Exception):
        b.run()  # no initial state
    sim = b.circuit(0)
    sim.h(0, 0)
    with pytest.raises(Exception):
        sim.m(0)  # readouts are refused on a recording context
    with pytest.raises(Exception):
        sim.getState()
    with pytest.raises(ValueError):
        sim.h(0, 3)  # invalid qubit: same validation as a normal context
    b.initFromSeparable([_init_arrays(np.random.default_rng(0), s) for s in ([1, 8], [2, 9])])
    b.run()
    with pytest.raises(ValueError):
        b.measure(2)  # no register 2
    with pytest.raises(ValueError):
        b.setLaunch(1024, 0)
    p = b.measure(0)
    assert p[0].shape == (2,) and p[1].shape == (4,)
    assert abs(p[0].sum() - 1.0) < 1e-5
