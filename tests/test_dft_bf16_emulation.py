"""The split-bf16 DFT-as-GEMM accuracy emulation of bench.py --dft-bf16-error:
bfloat16 rounding (nearest, ties to even) on known values, and the three-pass
error in the range the dropped lo*lo term and the bf16 lo parts set (~2^-18
relative per product; one dropped pass would show 2^-9)."""

import numpy as np

import bench


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -3.0 - 2 ** -7],
                 dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -3.0], dtype=np.float32)
    np.testing.assert_array_equal(bench._bf16(v), want)


def test_split_bf16_dft_error_range():
    r = bench.dft_bf16_split_error(Ns=(64,), cols=16)[0]
    assert 1e-6 < r["rel_l2_median"] < 2e-5
