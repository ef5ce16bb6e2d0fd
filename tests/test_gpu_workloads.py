"""The measured workloads at oracle-sized shapes: the state-transfer circuit
(bench.py's generator at [1]*6 + [10]) through the C ABI against the oracle:
final state, the six single-qubit marginals and the joint six-qubit marginal."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_state_transfer_marginals(prec):
    import bench
    from hipcvdv import HipCvdv, SeparableState
    from oracle.cvdv_oracle import CvdvOracle
    from oracle.separable_np import SeparableStateNP

    print(f"[workload] state transfer {prec}", flush=True)
    regs, cat, ops = bench.make_c3(nq=6, ncv=10)
    nq = len(regs) - 1
    h = HipCvdv(regs, precision=prec)
    o = CvdvOracle(regs, convention="cuda")
    sep, sepn = SeparableState(regs), SeparableStateNP(regs)
    for k in range(nq):
        sep.setZero(k)
        sepn.setZero(k)
    sep.setCat(nq, cat)
    sepn.setCat(nq, cat)
    h.initStateVector(sep)
    o.initStateVector(sepn)
    for op in ops:
        getattr(h, op[0])(*op[1:])
        getattr(o, op[0])(*op[1:])
    tol = 1e-12 if prec == "complex128" else 2e-6
    ref = o.getState().reshape(-1)
    assert np.max(np.abs(h.getState() - ref)) < tol * max(1.0, np.max(np.abs(ref)) * 10)
    for k in range(nq):
        np.testing.assert_allclose(h.m(k), o.m(k), atol=tol)
    joint = h.m(*range(nq))
    assert joint.shape == (2,) * nq
    np.testing.assert_allclose(joint, o.m(*range(nq)), atol=tol)
    assert abs(joint.sum() - 1.0) < (1e-12 if prec == "complex128" else 1e-5)


_C2_REF = {}


def _c2_oracle(regs):
    """The oracle (FP64, 'cuda' convention) after bench.make_circuit()'s 200 gates
    on the C2 initial state at a smaller register size (cached per shape)."""
    import bench
    from oracle.cvdv_oracle import CvdvOracle
    from oracle.separable_np import SeparableStateNP

    key = tuple(regs)
    if key not in _C2_REF:
        o = CvdvOracle(regs, convention="cuda")
        sep = SeparableStateNP(regs)
        sep.setCoherent(0, 1.2 - 0.4j)
        sep.setCoeffs(1, bench.squeezed_vacuum(regs[1], 0.5))
        o.initStateVector(sep)
        for op in bench.make_circuit():
            getattr(o, op[0])(*op[1:])
        _C2_REF[key] = o.getState().reshape(-1)
    return _C2_REF[key]


@pytest.mark.parametrize("regs", [[9, 9], [10, 10]], ids=lambda r: "x".join(map(str, r)))
@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_gaussian_circuit_against_oracle(prec, regs):
    """The headline workload's gate mix (bench.make_circuit: 200 seeded d, r, s,
    bs(pi/2), bs(theta); register 0 strided through the two-pass FFT, register 1
    contiguous) against the FP64 oracle: fidelity loss and max amplitude error.
    Bars: complex128 loss <= 1e-12 and error <= 1e-11 max|psi|; mixed loss <= 1e-6
    (the headline's bound) and error <= 2e-5 max|psi|."""
    import bench
    import torch
    from hipcvdv import HipCvdv, SeparableState

    print(f"[workload] gaussian circuit {prec} {regs}", flush=True)
    ref = _c2_oracle(regs)
    h = HipCvdv(regs, precision=prec)
    sep = SeparableState(regs)
    sep.setCoherent(0, 1.2 - 0.4j)
    sep.setCoeffs(1, torch.from_numpy(bench.squeezed_vacuum(regs[1], 0.5)))
    h.initStateVector(sep)
    for op in bench.make_circuit():
        getattr(h, op[0])(*op[1:])
    st = h.getState()
    nr, ns = np.linalg.norm(ref), np.linalg.norm(st)
    loss = 1.0 - abs(np.vdot(ref, st)) ** 2 / (nr * nr * ns * ns)
    err = np.max(np.abs(st - ref)) / np.max(np.abs(ref))
    print(f"[workload]   fidelity loss {loss:.3e}, max amp err / max|psi| {err:.3e}, "
          f"norm {ns * ns:.12f} (oracle {nr * nr:.12f})", flush=True)
    if prec == "complex128":
        assert loss <= 1e-12 and err <= 1e-11, (loss, err)
    else:
        assert loss <= 1e-6 and err <= 2e-5, (loss, err)
