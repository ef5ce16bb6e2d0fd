"""Pin the oracle against golden vectors written by the reference's TorchCvdv.

The fixtures (tests/golden/torchcvdv_golden.npz) were produced by
tests/golden/make_golden.py from src/torchCvdv.py and src/separable.py of the
reference, complex128 on the CPU.  Each case replays the same separable init,
gate sequence and readouts through the oracle (torch convention).
"""

import numpy as np
import pytest

from golden_cases import apply_ops, build_sep, c, load
from oracle.cvdv_oracle import CvdvOracle
from oracle.separable_np import SeparableStateNP

Z, CASES = load()


@pytest.mark.parametrize("case", CASES, ids=[cs["name"] for cs in CASES])
def test_oracle_matches_reference_golden(case):
    name = case["name"]
    sep = build_sep(SeparableStateNP, case["regs"], case["init"])
    for r, arr in enumerate(sep.register_arrays):
        np.testing.assert_allclose(arr, Z[f"{name}/sep{r}"], rtol=0, atol=1e-15)
    sim = CvdvOracle(case["regs"])
    sim.initStateVector(sep)
    np.testing.assert_allclose(sim.getState(), Z[f"{name}/init"], rtol=0, atol=1e-15)
    apply_ops(sim, case["ops"])
    np.testing.assert_allclose(sim.getState(), Z[f"{name}/final"], rtol=0, atol=1e-12)
    for i, rd in enumerate(case["reads"]):
        meth, *args = rd
        if meth == "m":
            got = sim.m(*args)
        elif meth == "getFidelity":
            got = sim.getFidelity(build_sep(SeparableStateNP, case["regs"], args[0]))
        else:
            got = getattr(sim, meth)(*args)
        np.testing.assert_allclose(np.asarray(got), Z[f"{name}/read{i}"], rtol=0, atol=1e-12,
                                   err_msg=f"{name} read {rd}")
