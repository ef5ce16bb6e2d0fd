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
