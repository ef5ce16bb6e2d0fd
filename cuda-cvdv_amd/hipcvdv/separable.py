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

    def _coherent(self, regIdx: int, alpha: complex) -> torch.Tensor:
        """pi^-1/4 sqrt(dx) e^{-(x-q)^2/2} e^{i(p x - p q/2)}, q = sqrt2 Re a, p = sqrt2 Im a."""
        x = self._x(regIdx)
        q, p = 2 ** 0.5 * alpha.real, 2 ** 0.5 * alpha.imag
        envelope = _PI_M_QUARTER * self.grid_steps[regIdx] ** 0.5 * torch.exp(-0.5 * (x - q) ** 2)
        return envelope * torch.exp(1j * (p * x - p * q / 2.0)).to(torch.complex128)

    def setCoherent(self, regIdx: int, alpha: Union[complex, float, int]) -> None:
        """Coherent state |alpha> on the grid (not renormalised)."""
        self.register_arrays[regIdx] = self._coherent(regIdx, complex(alpha))

    def _hermite(self, regIdx: int, count: int) -> torch.Tensor:
        """(count, N) real table of the grid-sampled Hermite functions
        h_n(x) pi^-1/4 sqrt(dx) e^{-x^2/2} (discrete norm ~1), n < count, by
        h_n = sqrt(2/n) x h_{n-1} - sqrt((n-1)/n) h_{n-2}."""
        x = self._x(regIdx)
        table = torch.empty((max(count, 1), x.numel()), dtype=torch.float64, device=self.device)
        table[0] = torch.exp(-0.5 * x * x) * _PI_M_QUARTER * self.grid_steps[regIdx] ** 0.5
        lower = torch.zeros_like(x)
        for n in range(1, count):
            table[n] = (2.0 / n) ** 0.5 * x * table[n - 1] - ((n - 1.0) / n) ** 0.5 * lower
            lower = table[n - 1]
        return table

    def setFock(self, regIdx: int, n: int) -> None:
        """Fock state |n> (grid-sampled Hermite function; not renormalised)."""
        self.register_arrays[regIdx] = self._hermite(regIdx, n + 1)[n].to(torch.complex128)

    def setUniform(self, regIdx: int) -> None:
        """Uniform superposition N^-1/2 sum_k |k>."""
        dim = self.register_dims[regIdx]
        self.register_arrays[regIdx] = torch.full((dim,), dim ** -0.5, dtype=torch.complex128,
                                                  device=self.device)

    def setFocks(self, regIdx: int, coeffs: Union[Sequence[complex], torch.Tensor]) -> None:
        """sum_n c_n |n> (not renormalised); all-zero coefficients give setZero."""
        c = torch.as_tensor(coeffs if isinstance(coeffs, torch.Tensor) else list(coeffs),
                            dtype=torch.complex128).to(self.device)
        if not bool((c.abs() > 1e-12).any()):
            self.setZero(regIdx)
            return
        table = self._hermite(regIdx, c.numel()).to(torch.complex128)
        self.register_arrays[regIdx] = (c.unsqueeze(1) * table).sum(dim=0)

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
        """sum_k c_k |alpha_k>, scaled to unit norm over the grid."""
        parts = [(complex(c), self._coherent(regIdx, complex(a))) for a, c in cat_states]
        weights = torch.tensor([c for c, _ in parts], dtype=torch.complex128, device=self.device)
        waves = torch.stack([w for _, w in parts])
        psi = (weights.unsqueeze(1) * waves).sum(dim=0)
        self.register_arrays[regIdx] = psi / psi.abs().square().sum().sqrt()

    def validate(self) -> None:
        """Raise RuntimeError naming the first register that no setter has filled."""
        missing = [r for r, a in enumerate(self.register_arrays) if a is None]
        if missing:
            raise RuntimeError(f"SeparableState: register {missing[0]} has no wavefunction "
                               f"(call one of the set* methods for it)")
