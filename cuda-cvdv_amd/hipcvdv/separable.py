"""SeparableState - per-register initial wavefunctions (product state).

Same builders, arguments and conventions as the reference's
``src/separable.py:12-138``: register arrays are 1-D complex128 torch tensors
(by default on the GPU, so ``HipCvdv.initStateVector`` can hand their device
pointers straight to ``cvdvInitFromSeparable``).  ``setFock``/``setFocks`` are
not renormalised and ``setCat`` is; ``setZero`` is the grid delta at j = 0.
"""

from __future__ import annotations

from typing import List, Optional, Sequence, Tuple, Union

import torch

_PI_M_QUARTER = torch.pi ** (-0.25)


class SeparableState:
    """Per-register product state container (separable.py:12-138)."""

    def __init__(self, numQubits_list: List[int], device: str = "cuda") -> None:
        self.device = torch.device(device)
        self.num_registers = len(numQubits_list)
        self.qubit_counts = list(numQubits_list)
        self.register_dims = [1 << n for n in numQubits_list]
        self.grid_steps = [(2 * torch.pi / d) ** 0.5 for d in self.register_dims]
        self.register_arrays: List[Optional[torch.Tensor]] = [None] * self.num_registers

    def _x(self, regIdx: int) -> torch.Tensor:
        dim = self.register_dims[regIdx]
        idx = torch.arange(dim, device=self.device, dtype=torch.float64)
        return (idx - (dim - 1) * 0.5) * self.grid_steps[regIdx]

    def setZero(self, regIdx: int) -> None:
        """Grid basis state |j = 0> (qubit |0>; not the oscillator vacuum)."""
        arr = torch.zeros(self.register_dims[regIdx], dtype=torch.complex128, device=self.device)
        arr[0] = 1.0
        self.register_arrays[regIdx] = arr

    def setCoherent(self, regIdx: int, alpha: Union[complex, float, int]) -> None:
        """Coherent state |alpha>: pi^-1/4 sqrt(dx) e^{-(x-q)^2/2} e^{i(p x - p q/2)}."""
        alpha = complex(alpha)
        x = self._x(regIdx)
        q = 2 ** 0.5 * alpha.real
        p = 2 ** 0.5 * alpha.imag
        amp = _PI_M_QUARTER * self.grid_steps[regIdx] ** 0.5 * torch.exp(-0.5 * (x - q) ** 2)
        self.register_arrays[regIdx] = amp * torch.exp(1j * (p * x - p * q / 2.0)).to(torch.complex128)

    def setFock(self, regIdx: int, n: int) -> None:
        """Fock state |n> by the Hermite-function three-term recurrence (not renormalised)."""
        x = self._x(regIdx)
        prev = torch.zeros_like(x)
        cur = torch.exp(-0.5 * x * x) * _PI_M_QUARTER * self.grid_steps[regIdx] ** 0.5
        for k in range(1, n + 1):
            prev, cur = cur, (2.0 / k) ** 0.5 * x * cur - ((k - 1.0) / k) ** 0.5 * prev
        self.register_arrays[regIdx] = cur.to(torch.complex128)

    def setUniform(self, regIdx: int) -> None:
        """Uniform superposition N^-1/2 sum_k |k>."""
        dim = self.register_dims[regIdx]
        self.register_arrays[regIdx] = torch.full((dim,), dim ** -0.5, dtype=torch.complex128,
                                                  device=self.device)

    def setFocks(self, regIdx: int, coeffs: Union[Sequence[complex], torch.Tensor]) -> None:
        """sum_n c_n |n> (not renormalised); all-zero coefficients give setZero."""
        c = torch.as_tensor(coeffs if isinstance(coeffs, torch.Tensor) else list(coeffs),
                            dtype=torch.complex128).to(self.device)
        if not bool(torch.any(torch.abs(c) > 1e-12)):
            self.setZero(regIdx)
            return
        x = self._x(regIdx)
        nmax = c.numel()
        basis = torch.zeros((nmax, x.numel()), dtype=torch.float64, device=self.device)
        basis[0] = torch.exp(-0.5 * x * x) * _PI_M_QUARTER * self.grid_steps[regIdx] ** 0.5
        if nmax > 1:
            basis[1] = 2.0 ** 0.5 * x * basis[0]
            for k in range(2, nmax):
                basis[k] = (2.0 / k) ** 0.5 * x * basis[k - 1] - ((k - 1.0) / k) ** 0.5 * basis[k - 2]
        self.register_arrays[regIdx] = torch.sum(c[:, None] * basis.to(torch.complex128), dim=0)

    def setCoeffs(self, regIdx: int, coeffs: Union[Sequence[complex], torch.Tensor]) -> None:
        """Arbitrary position-basis amplitudes (used as given)."""
        if isinstance(coeffs, torch.Tensor):
            arr = coeffs.to(dtype=torch.complex128, device=self.device)
        else:
            arr = torch.tensor(list(coeffs), dtype=torch.complex128, device=self.device)
        if arr.numel() != self.register_dims[regIdx]:
            raise ValueError(f"register {regIdx} needs {self.register_dims[regIdx]} coefficients, "
                             f"got {arr.numel()}")
        self.register_arrays[regIdx] = arr

    def setCat(self, regIdx: int,
               cat_states: Sequence[Tuple[Union[complex, float], Union[complex, float]]]) -> None:
        """Normalised superposition sum_k c_k |alpha_k> of coherent states."""
        x = self._x(regIdx)
        alphas = [complex(a) for a, _ in cat_states]
        coeffs = torch.tensor([complex(c) for _, c in cat_states], dtype=torch.complex128,
                              device=self.device)
        q = torch.tensor([2 ** 0.5 * a.real for a in alphas], dtype=torch.float64, device=self.device)
        p = torch.tensor([2 ** 0.5 * a.imag for a in alphas], dtype=torch.float64, device=self.device)
        gauss = torch.exp(-0.5 * (x[None, :] - q[:, None]) ** 2)
        phase = p[:, None] * x[None, :] - (p * q)[:, None] / 2.0
        coh = _PI_M_QUARTER * self.grid_steps[regIdx] ** 0.5 * gauss * torch.exp(1j * phase).to(torch.complex128)
        state = torch.sum(coeffs[:, None] * coh, dim=0)
        self.register_arrays[regIdx] = state / torch.sqrt(torch.sum(torch.abs(state) ** 2))

    def validate(self) -> None:
        """Raise if any register has not been initialised."""
        for i, arr in enumerate(self.register_arrays):
            if arr is None:
                raise RuntimeError(f"Register {i} not initialized in SeparableState")
