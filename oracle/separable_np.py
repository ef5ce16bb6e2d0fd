"""Oracle: numpy FP64 restatement of the reference's SeparableState builders.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``hipcvdv`` package,
``libcvdv_hip.so``) may import, call or link this module; only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use it, and
there only as the checker or as the timed CPU baseline.

Each builder follows ``src/separable.py`` of the reference (line numbers cited
per method).  Arrays are 1-D ``complex128`` numpy vectors, one per register.

Pinning (what checks each builder, besides the golden vectors written from the
reference's own SeparableState by ``tests/golden/make_golden.py``):
  * setCoherent / setFock / setFocks: discrete norm -> 1 for interior states,
    Hermite-function orthonormality, and the closed-form overlap
    |<alpha|beta>|^2 = exp(-|alpha-beta|^2)  (tests/test_oracle_closed_form.py).
  * setCat: normalised by construction; checked against a hand sum of
    coherent vectors.
  * setZero / setUniform / setCoeffs: trivial, checked directly.
"""

from __future__ import annotations

from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

PI_M_QUARTER = np.pi ** (-0.25)


class SeparableStateNP:
    """Per-register product state (numpy).  Mirrors separable.py:12-138."""

    def __init__(self, numQubits_list: List[int]) -> None:
        # separable.py:26-32
        self.num_registers = len(numQubits_list)
        self.qubit_counts = list(numQubits_list)
        self.register_dims = [1 << n for n in numQubits_list]
        self.grid_steps = [(2 * np.pi / d) ** 0.5 for d in self.register_dims]
        self.register_arrays: List[Optional[np.ndarray]] = [None] * self.num_registers

    def _x(self, regIdx: int) -> np.ndarray:
        # separable.py:36-40: x_j = (j - (N-1)/2) dx
        dim = self.register_dims[regIdx]
        dx = self.grid_steps[regIdx]
        return (np.arange(dim, dtype=np.float64) - (dim - 1) * 0.5) * dx

    def setZero(self, regIdx: int) -> None:
        # separable.py:44-48: grid delta at j=0 (not the vacuum)
        arr = np.zeros(self.register_dims[regIdx], dtype=np.complex128)
        arr[0] = 1.0
        self.register_arrays[regIdx] = arr

    def setCoherent(self, regIdx: int, alpha: Union[complex, float, int]) -> None:
        # separable.py:50-61
        alpha = complex(alpha)
        dx = self.grid_steps[regIdx]
        x = self._x(regIdx)
        q = 2 ** 0.5 * alpha.real
        p = 2 ** 0.5 * alpha.imag
        gauss = np.exp(-0.5 * (x - q) ** 2)
        phase = np.exp(1j * (p * x - p * q / 2.0))
        self.register_arrays[regIdx] = (PI_M_QUARTER * dx ** 0.5) * gauss * phase

    def setFock(self, regIdx: int, n: int) -> None:
        # separable.py:63-73: Hermite-function recurrence, not renormalised
        dx = self.grid_steps[regIdx]
        x = self._x(regIdx)
        prev = np.zeros_like(x)
        cur = np.exp(-0.5 * x * x) * PI_M_QUARTER * dx ** 0.5
        for k in range(1, n + 1):
            nxt = (2.0 / k) ** 0.5 * x * cur - ((k - 1.0) / k) ** 0.5 * prev
            prev, cur = cur, nxt
        self.register_arrays[regIdx] = cur.astype(np.complex128)

    def setUniform(self, regIdx: int) -> None:
        # separable.py:75-79
        dim = self.register_dims[regIdx]
        self.register_arrays[regIdx] = np.ones(dim, dtype=np.complex128) / dim ** 0.5

    def setFocks(self, regIdx: int, coeffs: Sequence[complex]) -> None:
        # separable.py:81-102: sum_n c_n |n>, not renormalised
        c = np.asarray(list(coeffs), dtype=np.complex128)
        if not np.any(np.abs(c) > 1e-12):
            self.setZero(regIdx)
            return
        dim = self.register_dims[regIdx]
        dx = self.grid_steps[regIdx]
        x = self._x(regIdx)
        nmax = len(c)
        fock = np.zeros((nmax, dim), dtype=np.float64)
        fock[0] = np.exp(-0.5 * x * x) * PI_M_QUARTER * dx ** 0.5
        if nmax > 1:
            fock[1] = 2.0 ** 0.5 * x * fock[0]
            for k in range(2, nmax):
                fock[k] = (2.0 / k) ** 0.5 * x * fock[k - 1] - ((k - 1.0) / k) ** 0.5 * fock[k - 2]
        self.register_arrays[regIdx] = np.sum(c[:, None] * fock.astype(np.complex128), axis=0)

    def setCoeffs(self, regIdx: int, coeffs: Sequence[complex]) -> None:
        # separable.py:104-111
        self.register_arrays[regIdx] = np.asarray(coeffs, dtype=np.complex128).copy()

    def setCat(self, regIdx: int,
               cat_states: Sequence[Tuple[Union[complex, float], Union[complex, float]]]) -> None:
        # separable.py:113-130: normalised sum of coherent states
        dx = self.grid_steps[regIdx]
        x = self._x(regIdx)
        alphas = [complex(a) for a, _ in cat_states]
        coeffs = np.asarray([complex(c) for _, c in cat_states], dtype=np.complex128)
        q = np.asarray([2 ** 0.5 * a.real for a in alphas])
        p = np.asarray([2 ** 0.5 * a.imag for a in alphas])
        gauss = np.exp(-0.5 * (x[None, :] - q[:, None]) ** 2)
        phase = p[:, None] * x[None, :] - (p * q)[:, None] / 2.0
        coh = PI_M_QUARTER * dx ** 0.5 * gauss * np.exp(1j * phase)
        state = np.sum(coeffs[:, None] * coh, axis=0)
        self.register_arrays[regIdx] = state / np.sqrt(np.sum(np.abs(state) ** 2))

    def validate(self) -> None:
        # separable.py:134-138
        for i, arr in enumerate(self.register_arrays):
            if arr is None:
                raise RuntimeError(f"Register {i} not initialized in SeparableState")
