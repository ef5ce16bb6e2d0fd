"""The product's SeparableState (torch, here on the CPU) against the reference's
own SeparableState outputs stored in the golden fixtures."""

import numpy as np
import pytest

from golden_cases import build_sep, load
from hipcvdv import SeparableState

Z, CASES = load()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_separable_builders_match_reference(case):
    sep = build_sep(SeparableState, case["regs"], case["init"], device="cpu")
    for r, arr in enumerate(sep.register_arrays):
        np.testing.assert_allclose(arr.numpy(), Z[f"{case['name']}/sep{r}"], rtol=0, atol=1e-15)


def test_validate_and_bad_coeffs():
    s = SeparableState([2, 3], device="cpu")
    s.setZero(0)
    with pytest.raises(RuntimeError):
        s.validate()
    with pytest.raises(ValueError):
        s.setCoeffs(1, [1.0, 0.0])
