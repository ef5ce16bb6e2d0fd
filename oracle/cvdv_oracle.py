"""Oracle: numpy/scipy FP64 restatement of the reference's TorchCvdv method.

TEST INFRASTRUCTURE ONLY.  The product (``hipcvdv`` + ``libcvdv_hip.so``) never
imports, calls or links anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use this
module: as the checker, or as the timed CPU baseline ("port").

Every method restates ``src/torchCvdv.py`` of the reference (file:line cited per
method) on a dense numpy array of shape ``(d_0, ..., d_{R-1})``, register 0 the
most significant axis (torchCvdv.py:26-27), in complex128.

Conventions that differ between the reference's two backends are selectable
(``convention='torch'`` is the default and the TorchCvdv behaviour; ``'cuda'``
restates the CudaCvdv/C-ABI behaviour the HIP library mirrors):

  =====================  ==============================  ===============================
  behaviour              torch (torchCvdv.py)            cuda (src/cuda/*.cu)
  =====================  ==============================  ===============================
  init normalisation     renormalise (:47-48)            none (context.cu:75-115)
  m() with no register   sqrt(sum |psi|^2) (:306-307)    sum |psi|^2 (measure.cu:276-300)
  r() turn count         np.round, half-even (:156)      rint, half-even (gates.cu:327)
  cr() turn count        floor(x+0.5) (:72)              rint, half-even (gates.cu:366)
  bs()/cbs() turn count  floor(x+0.5) (:204,231)         round, half-away (gates.cu:476,517)
  =====================  ==============================  ===============================

Pinning.  Every method is checked by at least one of: golden vectors written by
the reference's own TorchCvdv (tests/golden/make_golden.py ->
tests/golden/torchcvdv_golden.npz, checked in tests/test_oracle_golden.py), the
closed forms and identities of the reference's tests (tests/test_oracle_closed_form.py:
D(a)|0>=|a>, D(a)D(b)=D(a+b), FT|n>=(-i)^n|n>, CD commutator, R(t)|a>=|a e^{-it}>,
S conjugation, BS binomial, CR/CS/CBS identities, Wigner and Husimi closed forms,
<n>=|a|^2), and brute-force DFT/permutation checks.  No method is left
"parity unpinned".
"""

from __future__ import annotations

from typing import Any, List, Sequence

import numpy as np
import scipy.fft as sfft

from .separable_np import SeparableStateNP

PI = np.pi
SQRT2 = np.sqrt(2.0)


def _round_half_even(x: float) -> int:
    return int(np.rint(x))


def _round_half_away(x: float) -> int:
    return int(np.sign(x) * np.floor(abs(x) + 0.5))


def _floor_half_up(x: float) -> int:
    return int(np.floor(x + 0.5))


def _grid_split(u: float, dim: int):
    """u = position in grid units from x_0 -> (i, t), u = i + t, t in [0, 1); points
    within 1e-9 of a grid point snap to it (t = 0) so native points are exact."""
    r = round(u)
    if abs(u - r) < 1e-9:
        i, t = int(r), 0.0
    else:
        i = int(np.floor(u))
        t = u - i
    if i < 0:
        i, t = 0, 0.0
    if i > dim - 1:
        i, t = dim - 1, 0.0
    return i, t


class CvdvOracle:
    """FP64 dense reference simulator (restates torchCvdv.py:14-528)."""

    def __init__(self, numQubits_list: Sequence[int], convention: str = "torch",
                 workers: int = 1) -> None:
        if convention not in ("torch", "cuda"):
            raise ValueError("convention must be 'torch' or 'cuda'")
        # torchCvdv.py:17-28
        self.convention = convention
        self.workers = workers
        self.num_registers = len(numQubits_list)
        self.qubit_counts = np.array(list(numQubits_list), dtype=np.int32)
        self.register_dims = [1 << int(q) for q in numQubits_list]
        self.total_size = int(np.prod(self.register_dims, dtype=np.int64))
        self.grid_steps = np.array([np.sqrt(2 * PI / d) for d in self.register_dims])
        self.flwQbts = [int(sum(self.qubit_counts[r + 1:])) for r in range(self.num_registers)]
        self.state: Any = None

    # ------------------------------------------------------------------ helpers
    def _x(self, r: int) -> np.ndarray:
        # torchCvdv.py:418-424
        dim = self.register_dims[r]
        return (np.arange(dim, dtype=np.float64) - (dim - 1) * 0.5) * self.grid_steps[r]

    def _shape_for(self, axes: Sequence[int]) -> List[int]:
        shape = [1] * self.num_registers
        for a in axes:
            shape[a] = self.register_dims[a]
        return shape

    def _apply_phase(self, r: int, phase: np.ndarray) -> None:
        # torchCvdv.py:426-429
        self.state *= phase.reshape(self._shape_for([r]))

    def _ctrl_bits(self, ctrlReg: int, ctrlQubit: int) -> np.ndarray:
        # torchCvdv.py:465-466: bit (n_c - 1 - q) of the control register's local index
        n = int(self.qubit_counts[ctrlReg])
        return (np.arange(self.register_dims[ctrlReg]) >> (n - 1 - ctrlQubit)) & 1

    def _cond_factor(self, target_axes: Sequence[int], values: np.ndarray, ctrlReg: int,
                     ctrlQubit: int) -> np.ndarray:
        """Broadcastable factor e^{+i v} where the control bit is 0 and e^{-i v} where it
        is 1 (torchCvdv.py:459-469: Z convention).  ``values`` is defined over
        ``target_axes``; the control axis may coincide with a target axis."""
        bits = self._ctrl_bits(ctrlReg, ctrlQubit)
        vshape = [1] * self.num_registers
        for i, a in enumerate(target_axes):
            vshape[a] = values.shape[i]
        v = values.reshape(vshape)
        bshape = [1] * self.num_registers
        bshape[ctrlReg] = bits.size
        return np.where(bits.reshape(bshape) == 0, np.exp(1j * v), np.exp(-1j * v))

    # ------------------------------------------------------------------ init
    def initStateVector(self, sep: SeparableStateNP) -> None:
        # torchCvdv.py:38-48 (cuda convention: context.cu:75-115, no renormalisation)
        sep.validate()
        arrays = [np.asarray(a, dtype=np.complex128) for a in sep.register_arrays]
        if self.num_registers == 1:
            state = arrays[0].copy()
        else:
            state = np.multiply.outer(arrays[0], arrays[1])
            for i in range(2, self.num_registers):
                state = state[..., None] * arrays[i].reshape(*([1] * i), -1)
        if self.convention == "torch":
            state = state / np.sqrt(np.sum(np.abs(state) ** 2))
        self.state = np.ascontiguousarray(state)

    def initFromArray(self, arr) -> None:
        # cudaCvdv.py:162-169 (flat complex128 state, copied)
        self.state = np.asarray(arr, dtype=np.complex128).reshape(self.register_dims).copy()

    # ------------------------------------------------------------------ CV gates
    def d(self, regIdx: int, beta) -> None:
        # torchCvdv.py:50-59
        beta = complex(beta)
        if abs(beta.imag) > 1e-12:
            x = self._x(regIdx)
            self._apply_phase(regIdx, np.exp(1j * SQRT2 * beta.imag * x))
        if abs(beta.real) > 1e-12:
            self.ftQ2P(regIdx)
            p = self._x(regIdx)
            self._apply_phase(regIdx, np.exp(-1j * SQRT2 * beta.real * p))
            self.ftP2Q(regIdx)

    def cd(self, targetReg: int, ctrlReg: int, ctrlQubit: int, alpha) -> None:
        # torchCvdv.py:61-68
        alpha = complex(alpha)
        if abs(alpha.imag) > 1e-12:
            self._cond_phase_q(targetReg, ctrlReg, ctrlQubit, SQRT2 * alpha.imag)
        if abs(alpha.real) > 1e-12:
            self.ftQ2P(targetReg)
            self._cond_phase_q(targetReg, ctrlReg, ctrlQubit, -SQRT2 * alpha.real)
            self.ftP2Q(targetReg)

    def _turns(self, ratio: float, kind: str) -> int:
        if self.convention == "torch":
            return _round_half_even(ratio) if kind == "r" else _floor_half_up(ratio)
        # cuda: rint for r/cr, round (half away from zero) for bs/cbs
        return _round_half_even(ratio) if kind in ("r", "cr") else _round_half_away(ratio)

    def cr(self, targetReg: int, ctrlReg: int, ctrlQubit: int, theta) -> None:
        # torchCvdv.py:70-93 (cuda: gates.cu:350-392, 598-611)
        k = self._turns(theta / (PI / 2), "cr")
        remainder = theta - k * (PI / 2)
        q = (k % 4 + 4) % 4
        if q == 1:
            self.ftQ2P(targetReg)
            self.cp(targetReg, ctrlReg, ctrlQubit)
            self.rz(ctrlReg, ctrlQubit, PI / 2)
        elif q == 2:
            self.p(targetReg)
            self.rz(ctrlReg, ctrlQubit, PI)
        elif q == 3:
            self.ftP2Q(targetReg)
            self.cp(targetReg, ctrlReg, ctrlQubit)
            self.rz(ctrlReg, ctrlQubit, -PI / 2)
        if abs(remainder) > 1e-15:
            th = np.tan(remainder / 2)
            st = np.sin(remainder)
            self._cond_phase_q2(targetReg, ctrlReg, ctrlQubit, -0.5 * th)
            self.ftQ2P(targetReg)
            self._cond_phase_q2(targetReg, ctrlReg, ctrlQubit, -0.5 * st)
            self.ftP2Q(targetReg)
            self._cond_phase_q2(targetReg, ctrlReg, ctrlQubit, -0.5 * th)

    def sheer(self, regIdx: int, t) -> None:
        # torchCvdv.py:146-148
        x = self._x(regIdx)
        self._apply_phase(regIdx, np.exp(1j * t * x ** 2))

    def phaseCubic(self, regIdx: int, t) -> None:
        # torchCvdv.py:150-152
        x = self._x(regIdx)
        self._apply_phase(regIdx, np.exp(1j * t * x ** 3))

    def r(self, regIdx: int, theta) -> None:
        # torchCvdv.py:154-170 (cuda: gates.cu:316-348, 572-584)
        k = self._turns(theta / (PI / 2), "r")
        remainder = theta - k * (PI / 2)
        q = (k % 4 + 4) % 4
        if q == 1:
            self.ftQ2P(regIdx)
        elif q == 2:
            self.p(regIdx)
        elif q == 3:
            self.ftP2Q(regIdx)
        if abs(remainder) > 1e-15:
            self.sheer(regIdx, -0.5 * np.tan(remainder / 2))
            self.ftQ2P(regIdx)
            self.sheer(regIdx, -0.5 * np.sin(remainder))
            self.ftP2Q(regIdx)
            self.sheer(regIdx, -0.5 * np.tan(remainder / 2))

    def s(self, regIdx: int, r) -> None:
        # torchCvdv.py:172-184 (cuda: gates.cu:394-421)
        ehr = np.exp(0.5 * r)
        emhr = np.exp(-0.5 * r)
        sq = np.sqrt(abs(np.exp(r) - 1.0))
        sign = 1 if r >= 0 else -1
        self.ftQ2P(regIdx)
        self.sheer(regIdx, -0.5 * ehr * sq)
        self.ftP2Q(regIdx)
        self.sheer(regIdx, 0.5 * sign * emhr * sq)
        self.ftQ2P(regIdx)
        self.sheer(regIdx, 0.5 * emhr * sq)
        self.ftP2Q(regIdx)
        self.sheer(regIdx, -0.5 * sign * ehr * sq)

    def cs(self, targetReg: int, ctrlReg: int, ctrlQubit: int, r) -> None:
        # torchCvdv.py:186-200 (cuda: gates.cu:423-459); r=0 gives NaN in both
        ch = np.cosh(r)
        sh = np.sinh(r)
        sv = np.sqrt(2.0 * abs(np.sinh(0.5 * r)))
        with np.errstate(divide="ignore", invalid="ignore"):
            self.sheer(targetReg, 0.5 * sv * ch)
            self._cond_phase_q2(targetReg, ctrlReg, ctrlQubit, -0.5 * sv * sh)
            self.ftQ2P(targetReg)
            self.sheer(targetReg, 0.5 * (ch - 1) / sv)
            self._cond_phase_q2(targetReg, ctrlReg, ctrlQubit, 0.5 * sh / sv)
            self.ftP2Q(targetReg)
            self.sheer(targetReg, -0.5 * sv)
            self.ftQ2P(targetReg)
            self.sheer(targetReg, 0.5 * (ch - 1) / sv)
            self._cond_phase_q2(targetReg, ctrlReg, ctrlQubit, -0.5 * sh / sv)
            self.ftP2Q(targetReg)

    def bs(self, reg1: int, reg2: int, theta) -> None:
        # torchCvdv.py:202-227 (cuda: gates.cu:461-502, 614-636)
        k = self._turns(theta / PI, "bs")
        remainder = theta - k * PI
        h = (k % 4 + 4) % 4
        if h == 1:
            self.ftQ2P(reg1); self.ftQ2P(reg2); self.swap(reg1, reg2)
        elif h == 2:
            self.p(reg1); self.p(reg2)
        elif h == 3:
            self.ftP2Q(reg1); self.ftP2Q(reg2); self.swap(reg1, reg2)
        if abs(remainder) > 1e-15:
            tq = np.tan(remainder / 4)
            sh = np.sin(remainder / 2)
            self.q1q2(reg1, reg2, -tq)
            self.ftQ2P(reg1); self.ftQ2P(reg2)
            self.q1q2(reg1, reg2, -sh)
            self.ftP2Q(reg1); self.ftP2Q(reg2)
            self.q1q2(reg1, reg2, -tq)

    def cbs(self, reg1: int, reg2: int, ctrlReg: int, ctrlQubit: int, theta) -> None:
        # torchCvdv.py:229-258 (cuda: gates.cu:504-547, 640-666)
        k = self._turns(theta / PI, "bs")
        remainder = theta - k * PI
        h = (k % 4 + 4) % 4
        if h == 1:
            self.ftQ2P(reg1); self.ftQ2P(reg2)
            self.cp(reg1, ctrlReg, ctrlQubit); self.cp(reg2, ctrlReg, ctrlQubit)
            self.swap(reg1, reg2)
        elif h == 2:
            self.p(reg1); self.p(reg2)
        elif h == 3:
            self.ftP2Q(reg1); self.ftP2Q(reg2)
            self.cp(reg1, ctrlReg, ctrlQubit); self.cp(reg2, ctrlReg, ctrlQubit)
            self.swap(reg1, reg2)
        if abs(remainder) > 1e-15:
            tq = np.tan(remainder / 4)
            sh = np.sin(remainder / 2)
            self._cond_q1q2(reg1, reg2, ctrlReg, ctrlQubit, -tq)
            self.ftQ2P(reg1); self.ftQ2P(reg2)
            self._cond_q1q2(reg1, reg2, ctrlReg, ctrlQubit, -sh)
            self.ftP2Q(reg1); self.ftP2Q(reg2)
            self._cond_q1q2(reg1, reg2, ctrlReg, ctrlQubit, -tq)

    def q1q2(self, reg1: int, reg2: int, coeff) -> None:
        # torchCvdv.py:260-267
        q1 = self._x(reg1)
        q2 = self._x(reg2)
        pm = np.exp(1j * coeff * q1[:, None] * q2[None, :])
        if reg1 < reg2:
            self.state *= pm.reshape(self._shape_for([reg1, reg2]))
        elif self.convention == "torch":
            # torchCvdv.py:264-267 reshapes (does not transpose) the (d_reg1, d_reg2)
            # matrix onto axes (reg2, reg1); for unequal sizes this scrambles the
            # phase.  Restated as is; the cuda convention (kernelPhaseXX,
            # gates.cuh:213-231) applies e^{i c q1 q2} per element.
            self.state *= pm.reshape(self._shape_for([reg1, reg2]))
        else:
            self.state *= pm.T.reshape(self._shape_for([reg2, reg1]))

    def ftQ2P(self, regIdx: int) -> None:
        # torchCvdv.py:269-280 (cuda: gates.cu:19-67)
        dx = self.grid_steps[regIdx]
        dim = self.register_dims[regIdx]
        c = PI * (dim - 1.0) / (dim * dx)
        x = self._x(regIdx)
        self._apply_phase(regIdx, np.exp(1j * c * x))
        self.state = sfft.fft(self.state, axis=regIdx, workers=self.workers)
        self._apply_phase(regIdx, np.exp(1j * c * x))
        self.state /= np.sqrt(dim)
        self.state *= np.exp(1j * PI * (dim - 1) ** 2 / (2 * dim))

    def ftP2Q(self, regIdx: int) -> None:
        # torchCvdv.py:282-293 (cuda: gates.cu:69-122)
        dx = self.grid_steps[regIdx]
        dim = self.register_dims[regIdx]
        c = -PI * (dim - 1.0) / (dim * dx)
        x = self._x(regIdx)
        self._apply_phase(regIdx, np.exp(1j * c * x))
        self.state = sfft.ifft(self.state, axis=regIdx, norm="forward", workers=self.workers)
        self._apply_phase(regIdx, np.exp(1j * c * x))
        self.state /= np.sqrt(dim)
        self.state *= np.exp(-1j * PI * (dim - 1) ** 2 / (2 * dim))

    # ------------------------------------------------------------------ DV gates
    def _qubit_gate(self, regIdx: int, q: int, m: np.ndarray) -> None:
        # torchCvdv.py:431-445: pair (bit=0, bit=1) at bit n_r-1-q of the local index
        n = int(self.qubit_counts[regIdx])
        d = self.register_dims
        shp = list(d[:regIdx]) + [1 << q, 2, 1 << (n - 1 - q)] + list(d[regIdx + 1:])
        v = self.state.reshape(shp)
        ax = regIdx + 1
        s0 = np.take(v, 0, axis=ax)
        s1 = np.take(v, 1, axis=ax)
        n0 = m[0, 0] * s0 + m[0, 1] * s1
        n1 = m[1, 0] * s0 + m[1, 1] * s1
        self.state = np.stack([n0, n1], axis=ax).reshape(d)

    def x(self, regIdx: int, targetQubit: int) -> None:
        self.rx(regIdx, targetQubit, PI)  # torchCvdv.py:95

    def y(self, regIdx: int, targetQubit: int) -> None:
        self.ry(regIdx, targetQubit, PI)  # torchCvdv.py:96

    def z(self, regIdx: int, targetQubit: int) -> None:
        self.rz(regIdx, targetQubit, PI)  # torchCvdv.py:97

    def rx(self, regIdx: int, targetQubit: int, theta) -> None:
        # torchCvdv.py:99-103
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        self._qubit_gate(regIdx, targetQubit, np.array([[c, -1j * s], [-1j * s, c]]))

    def ry(self, regIdx: int, targetQubit: int, theta) -> None:
        # torchCvdv.py:105-109
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        self._qubit_gate(regIdx, targetQubit, np.array([[c, -s], [s, c]], dtype=np.complex128))

    def rz(self, regIdx: int, targetQubit: int, theta) -> None:
        # torchCvdv.py:111-113
        m = np.array([[np.exp(-1j * theta / 2), 0], [0, np.exp(1j * theta / 2)]])
        self._qubit_gate(regIdx, targetQubit, m)

    def h(self, regIdx: int, targetQubit: int) -> None:
        # torchCvdv.py:115-117
        m = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
        self._qubit_gate(regIdx, targetQubit, m)

    def p(self, regIdx: int) -> None:
        # torchCvdv.py:119-120
        self.state = np.ascontiguousarray(np.flip(self.state, axis=regIdx))

    def cp(self, targetReg: int, ctrlReg: int, ctrlQubit: int) -> None:
        # torchCvdv.py:122-137: flat-index pair swap (j <-> N-1-j) where the
        # control bit of the lower partner is 1
        dim = self.register_dims[targetReg]
        if ctrlReg != targetReg:
            bits = self._ctrl_bits(ctrlReg, ctrlQubit)
            bshape = self._shape_for([ctrlReg])
            flipped = np.flip(self.state, axis=targetReg)
            self.state = np.ascontiguousarray(
                np.where(bits.reshape(bshape) == 1, flipped, self.state))
            return
        flat = self.state.reshape(-1).copy()
        idx = np.arange(self.total_size, dtype=np.int64)
        pos = self.flwQbts[ctrlReg] + int(self.qubit_counts[ctrlReg]) - 1 - ctrlQubit
        cbit = (idx >> pos) & 1
        loc = (idx >> self.flwQbts[targetReg]) & (dim - 1)
        fl = (dim - 1) - loc
        act = (cbit == 1) & (loc < fl)
        src = idx[act]
        dst = src + (fl[act] - loc[act]) * (1 << self.flwQbts[targetReg])
        old = self.state.reshape(-1)
        flat[src] = old[dst]
        flat[dst] = old[src]
        self.state = flat.reshape(self.register_dims)

    def swap(self, reg1: int, reg2: int) -> None:
        # torchCvdv.py:139-144
        if self.qubit_counts[reg1] != self.qubit_counts[reg2]:
            raise ValueError("SWAP requires registers with same number of qubits")
        perm = list(range(self.num_registers))
        perm[reg1], perm[reg2] = perm[reg2], perm[reg1]
        self.state = np.ascontiguousarray(np.transpose(self.state, perm))

    # ------------------------------------------------------------------ conditional phases
    def _cond_phase_q(self, t: int, c: int, q: int, coeff) -> None:
        # torchCvdv.py:447-451
        x = self._x(t)
        self.state *= self._cond_factor([t], coeff * x, c, q)

    def _cond_phase_q2(self, t: int, c: int, q: int, tt) -> None:
        # torchCvdv.py:453-457
        x = self._x(t)
        self.state *= self._cond_factor([t], tt * x ** 2, c, q)

    def _cond_q1q2(self, r1: int, r2: int, c: int, q: int, coeff) -> None:
        # torchCvdv.py:471-484 (phase_val in FP64 for complex128)
        idx = np.arange(self.total_size, dtype=np.int64)
        l1 = (idx >> self.flwQbts[r1]) & (self.register_dims[r1] - 1)
        l2 = (idx >> self.flwQbts[r2]) & (self.register_dims[r2] - 1)
        pos = self.flwQbts[c] + int(self.qubit_counts[c]) - 1 - q
        cb = (idx >> pos) & 1
        pv = coeff * self._x(r1)[l1] * self._x(r2)[l2]
        ph = np.where(cb == 0, np.exp(1j * pv), np.exp(-1j * pv))
        self.state = (self.state.reshape(-1) * ph).reshape(self.register_dims)

    # ------------------------------------------------------------------ readout
    def getState(self) -> np.ndarray:
        # torchCvdv.py:295-296
        return np.ascontiguousarray(self.state).reshape(-1).copy()

    def getXGrid(self, regIdx: int) -> np.ndarray:
        # torchCvdv.py:298-301
        dim = self.register_dims[regIdx]
        return (np.arange(dim) - (dim - 1) * 0.5) * self.grid_steps[regIdx]

    def m(self, *regIdxs):
        # torchCvdv.py:303-318 (cuda convention: no sqrt, measure.cu:276-300)
        if len(regIdxs) == 1 and isinstance(regIdxs[0], (list, tuple)):
            regIdxs = tuple(regIdxs[0])
        if not regIdxs:
            n2 = float(np.sum(np.abs(self.state) ** 2))
            return float(np.sqrt(n2)) if self.convention == "torch" else n2
        regs = list(regIdxs)
        probs = np.abs(self.state) ** 2
        for d in sorted([i for i in range(self.num_registers) if i not in regs], reverse=True):
            probs = np.sum(probs, axis=d)
        sreg = sorted(regs)
        perm = [sreg.index(r) for r in regs]
        if perm != list(range(len(regs))):
            probs = np.transpose(probs, perm)
        return np.ascontiguousarray(probs)

    def _product(self, sep: SeparableStateNP) -> np.ndarray:
        arrays = [np.asarray(a, dtype=np.complex128) for a in sep.register_arrays]
        if self.num_registers == 1:
            return arrays[0]
        ref = np.multiply.outer(arrays[0], arrays[1])
        for i in range(2, self.num_registers):
            ref = ref[..., None] * arrays[i].reshape(*([1] * i), -1)
        return ref

    def getFidelity(self, sep: SeparableStateNP) -> float:
        # torchCvdv.py:320-331
        sep.validate()
        ref = self._product(sep).reshape(-1)
        inner = np.sum(np.conj(ref) * self.state.reshape(-1))
        return float(abs(inner) ** 2)

    def fidelityWith(self, other: "CvdvOracle") -> float:
        # measure.cu:333-366: |<psi1|psi2>|^2
        inner = np.sum(np.conj(self.state.reshape(-1)) * other.state.reshape(-1))
        return float(abs(inner) ** 2)

    def getPhotonNumber(self, regIdx: int) -> float:
        # torchCvdv.py:333-342
        x = self._x(regIdx)
        shape = self._shape_for([regIdx])
        q2 = float(np.sum(np.abs(self.state) ** 2 * x.reshape(shape) ** 2))
        self.ftQ2P(regIdx)
        p2 = float(np.sum(np.abs(self.state) ** 2 * x.reshape(shape) ** 2))
        self.ftP2Q(regIdx)
        return (q2 + p2 - 1.0) / 2.0

    def _slices(self, regIdx: int) -> np.ndarray:
        # torchCvdv.py:406-413: (num_slices, dim)
        if self.num_registers == 1:
            return self.state.reshape(1, -1)
        perm = list(range(self.num_registers))
        perm[0], perm[regIdx] = perm[regIdx], perm[0]
        st = np.ascontiguousarray(np.transpose(self.state, perm))
        return st.reshape(self.register_dims[regIdx], -1).T

    def getWigner(self, regIdx: int, bound: float | None = None) -> np.ndarray:
        # torchCvdv.py:344-374 (includes the e^{-i p (N-1) dx} finalisation)
        dim = self.register_dims[regIdx]
        dx = self.grid_steps[regIdx]
        sl = self._slices(regIdx)
        sxs = sl.T
        rho = sxs @ np.conj(sxs).T
        i_idx = np.arange(dim)
        kd = np.arange(dim) - (dim - 1) // 2
        ipy = i_idx[:, None] + kd[None, :]
        imy = i_idx[:, None] - kd[None, :]
        valid = (ipy >= 0) & (ipy < dim) & (imy >= 0) & (imy < dim)
        integ = np.where(valid, rho[np.clip(imy, 0, dim - 1), np.clip(ipy, 0, dim - 1)], 0)
        fr = sfft.ifft(integ, axis=1, workers=self.workers) * dim
        dp = PI / (dim * dx)
        jc = np.arange(dim)
        kf = (jc + dim // 2) % dim
        pj = (jc - dim / 2) * dp
        pc = np.exp(-1j * (dim - 1) * dx * pj)
        W = (pc[:, None] * fr.T[kf]).real * (dx / PI)
        if bound is None:
            return W
        xv = (np.arange(dim) - (dim - 1) / 2) * dx
        pv = (np.arange(dim) - dim / 2) * dp
        return W[np.ix_(np.abs(pv) <= bound, np.abs(xv) <= bound)]

    def _slices_canonical(self, regIdx: int) -> np.ndarray:
        """(S, dim) slices of register regIdx, slice index = flat index of the other
        registers in their canonical order (register 0 most significant).  The
        traced readouts do not depend on the order; per-slice output does."""
        return np.moveaxis(self.state, regIdx, -1).reshape(-1, self.register_dims[regIdx])

    def getWignerGrid(self, regIdx: int, xs, ps, per_slice: bool = False,
                      native: bool = True) -> np.ndarray:
        """Extension (no reference counterpart): the discrete Wigner sum of getWigner
        (torchCvdv.py:344-374) at arbitrary points.  With x = x_i + t dx (t in [0, 1),
        t = 0 on grid points) and A(d) = sum_s phi_s(i-d) conj(phi_s(i+d)), phi_s the
        trigonometric interpolant of slice s sampled at x_j + t dx:
            W(x, p) = dx/pi Re[e^{-i (N-1) dx p} sum_d A(d) e^{2 i p dx (d + (N-1)//2)}]
        (native=True; on the native grid this is getWigner exactly), or without the
        constant-offset phase (native=False): dx/pi Re[sum_d A(d) e^{2 i p dx d}].
        Output (len(ps), len(xs)); per_slice: (S, len(ps), len(xs)), slices in the
        canonical order of the other registers."""
        dim = self.register_dims[regIdx]
        dx = self.grid_steps[regIdx]
        sl = self._slices_canonical(regIdx)
        S = sl.shape[0]
        xs = np.asarray(xs, dtype=np.float64)
        ps = np.asarray(ps, dtype=np.float64)
        c = np.fft.fft(sl, axis=1)
        kk = np.fft.fftfreq(dim) * dim
        x0 = -(dim - 1) / 2 * dx
        o = (dim - 1) // 2
        out = np.zeros((S if per_slice else 1, len(ps), len(xs)))
        for ix, x in enumerate(xs):
            i, t = _grid_split((x - x0) / dx, dim)
            phi = sl if t == 0.0 else np.fft.ifft(c * np.exp(2j * PI * kk * t / dim), axis=1)
            dmax = min(i, dim - 1 - i)
            d = np.arange(-dmax, dmax + 1)
            A = phi[:, i - d] * np.conj(phi[:, i + d])  # (S, nd)
            if not per_slice:
                A = A.sum(axis=0, keepdims=True)
            if native:
                E = np.exp(1j * dx * (np.outer(ps, 2 * (d + o)) - (dim - 1) * ps[:, None]))
            else:
                E = np.exp(2j * dx * np.outer(ps, d))
            out[:, :, ix] = (A @ E.T).real * (dx / PI)
        return out if per_slice else out[0]

    def getHusimiQGrid(self, regIdx: int, qs, ps, per_slice: bool = False) -> np.ndarray:
        """Extension (no reference counterpart): getHusimiQ's sum (torchCvdv.py:376-404)
        at arbitrary points, Q(q, p) = dx/pi^(3/2) sum_s |sum_j psi_s(x_j)
        e^{-(x_j-q)^2/2} e^{-i p x_j}|^2 (on the native grid this is getHusimiQ
        exactly).  Output (len(ps), len(qs)); per_slice: (S, len(ps), len(qs))."""
        dx = self.grid_steps[regIdx]
        sl = self._slices_canonical(regIdx)
        xg = self._x(regIdx)
        qs = np.asarray(qs, dtype=np.float64)
        ps = np.asarray(ps, dtype=np.float64)
        B = sl[:, :, None] * np.exp(-0.5 * (xg[:, None] - qs[None, :]) ** 2)[None]  # (S, N, nq)
        E = np.exp(-1j * np.outer(ps, xg))  # (np, N)
        H = np.einsum("pj,sjq->spq", E, B)
        Q = np.abs(H) ** 2 * (dx / PI ** 1.5)
        return Q if per_slice else Q.sum(axis=0)

    def getHusimiQWigner(self, regIdx: int) -> np.ndarray:
        """The CUDA backend's fallback Husimi route (bosonic.cu:63-101, kernels
        bosonic.cuh:158-186), which TorchCvdv lacks: the native Wigner (getWigner,
        rows p with spacing pi/(N dx), columns x), its 2D DFT times
        g = exp(-(wq^2 + wp^2)/4) with wq = 2 pi kq/(N dx) along x and
        wp = 2 pi kp/(N dp) along p (signed frequencies), the inverse 2D DFT, its
        real part clamped at 0.  A circular convolution of W with a Gaussian of
        variance 1/2 per axis: on the Wigner's grid, not getHusimiQ's, and
        normalised like W (a coherent state gives dx/(2 pi) e^{-|z - z0|^2/2})."""
        dim = self.register_dims[regIdx]
        dx = self.grid_steps[regIdx]
        W = self.getWigner(regIdx)
        k = np.fft.fftfreq(dim) * dim  # signed index: k for k <= N/2, k - N above
        k[dim // 2] = dim // 2          # bosonic.cuh:166-167 keeps +N/2 positive
        dp = PI / (dim * dx)
        wq = 2 * PI * k / (dim * dx)
        wp = 2 * PI * k / (dim * dp)
        g = np.exp(-0.25 * (wp[:, None] ** 2 + wq[None, :] ** 2))
        Q = np.fft.ifft2(np.fft.fft2(W) * g).real
        return np.maximum(Q, 0.0)

    def getHusimiQ(self, regIdx: int, bound: float | None = None) -> np.ndarray:
        # torchCvdv.py:376-404 (non-periodic Gaussian window)
        dim = self.register_dims[regIdx]
        dx = self.grid_steps[regIdx]
        sl = self._slices(regIdx)
        xg = self._x(regIdx)
        window = PI ** -0.25 * np.exp(-0.5 * (xg[None, :] - xg[:, None]) ** 2) * np.sqrt(dx)
        acc = np.zeros((dim, dim))
        for s in range(sl.shape[0]):
            H = sfft.fft(sl[s][None, :] * window, axis=1, workers=self.workers)
            acc += np.abs(H) ** 2
        kf = (np.arange(dim) + dim // 2) % dim
        Q = acc[:, kf].T / PI
        if bound is None:
            return Q
        dp = 2 * PI / (dim * dx)
        qv = (np.arange(dim) - (dim - 1) / 2) * dx
        pv = (np.arange(dim) - dim / 2) * dp
        return Q[np.ix_(np.abs(pv) <= bound, np.abs(qv) <= bound)]
