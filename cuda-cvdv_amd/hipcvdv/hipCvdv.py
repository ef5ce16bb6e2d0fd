"""HipCvdv - drop-in counterpart of the reference's ``CudaCvdv`` (src/cudaCvdv.py:77-252).

Same constructor, method names, argument order, defaults and return shapes;
every call goes through the ctypes C ABI of ``libcvdv_hip.so``.  Additions:
the ``precision`` keyword ('complex128' as in the reference, or 'mixed' for
complex64 amplitudes with FP64 phase angles, twiddles and sums), and
``synchronize()``, ``flush()`` and ``info()``.

Errors recorded by the library are raised (ValueError for invalid arguments,
CvdvError for runtime failures); the reference prints them and continues.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int
from typing import Any, List, Tuple

import numpy as np
import numpy.typing as npt
import torch

from . import _lib
from .separable import SeparableState

PRECISIONS = {"complex128": 0, "c128": 0, "fp64": 0, "mixed": 1, "complex64": 1, "c64": 1}


def _dptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class HipCvdv:
    """Dense hybrid CV-DV state-vector simulator on an MI355X (gfx950)."""

    def __init__(self, numQubits_list: List[int], precision: str = "complex128") -> None:
        # cudaCvdv.py:80-87
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.num_registers = len(numQubits_list)
        self.register_dims = [1 << q for q in numQubits_list]
        self.total_size = int(np.prod(self.register_dims, dtype=np.int64))
        self.qubit_counts = np.array(numQubits_list, dtype=np.int32)
        self.grid_steps = np.array([np.sqrt(2 * np.pi / d) for d in self.register_dims], dtype=np.float64)
        self.precision = "complex128" if PRECISIONS[precision] == 0 else "mixed"
        self._lib = _lib.get_lib()
        arr = (c_int * len(numQubits_list))(*numQubits_list)
        self.ctx = self._lib.cvdvCreateEx(len(numQubits_list), arr, PRECISIONS[precision])
        _lib.check()

    @classmethod
    def _recording(cls, lib, ctx: int, numQubits_list: List[int], precision: str, owner: Any) -> "HipCvdv":
        """Wrap a batch circuit's recording context (gates only; owned by ``owner``)."""
        self = cls.__new__(cls)
        self.num_registers = len(numQubits_list)
        self.register_dims = [1 << q for q in numQubits_list]
        self.total_size = int(np.prod(self.register_dims, dtype=np.int64))
        self.qubit_counts = np.array(numQubits_list, dtype=np.int32)
        self.grid_steps = np.array([np.sqrt(2 * np.pi / d) for d in self.register_dims], dtype=np.float64)
        self.precision = precision
        self._lib = lib
        self.ctx = ctx
        self._owner = owner  # keeps the batch (and so the context) alive
        return self

    def __del__(self):
        try:
            if getattr(self, "_owner", None) is not None:
                return
            if getattr(self, "ctx", None):
                self._lib.cvdvDestroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    def _device_arrays(self, sep: SeparableState):
        sep.validate()
        if len(sep.register_arrays) != self.num_registers:
            raise ValueError(f"SeparableState has {len(sep.register_arrays)} registers, "
                             f"simulator has {self.num_registers}")
        arrays = []
        for r, a in enumerate(sep.register_arrays):
            if a.numel() != self.register_dims[r]:
                raise ValueError(f"register {r}: {a.numel()} amplitudes, expected {self.register_dims[r]}")
            arrays.append(a.to(device="cuda", dtype=torch.complex128).contiguous())
        return arrays

    def _c(self, name: str, *args):
        r = getattr(self._lib, name)(self.ctx, *args)
        _lib.check()
        return r

    # ------------------------------------------------------------ init
    def initStateVector(self, sep: SeparableState) -> None:
        """Tensor product of the register arrays (not renormalised, context.cu:75-115)."""
        arrays = self._device_arrays(sep)
        ptrs = (ctypes.c_void_p * self.num_registers)(*[_dptr(a) for a in arrays])
        self._c("cvdvInitFromSeparable", ptrs, c_int(self.num_registers))

    def initFromArray(self, arr) -> None:
        """Load a full flat complex128 state (cudaCvdv.py:162-169)."""
        if isinstance(arr, torch.Tensor):
            t = arr.reshape(-1).to(dtype=torch.complex128, device="cuda").contiguous()
        else:
            t = torch.from_numpy(np.asarray(arr, dtype=np.complex128).reshape(-1)).to("cuda").contiguous()
        if t.numel() != self.total_size:
            raise ValueError(f"state has {t.numel()} amplitudes, expected {self.total_size}")
        self._c("cvdvSetStateFromDevicePtr", _dptr(t))

    # ------------------------------------------------------------ gates (cudaCvdv.py:102-132)
    def d(self, regIdx: int, beta) -> None:
        beta = complex(beta)
        self._c("cvdvDisplacement", regIdx, c_double(beta.real), c_double(beta.imag))

    def cd(self, targetReg: int, ctrlReg: int, ctrlQubit: int, alpha) -> None:
        alpha = complex(alpha)
        self._c("cvdvConditionalDisplacement", targetReg, ctrlReg, ctrlQubit,
                c_double(alpha.real), c_double(alpha.imag))

    def cr(self, targetReg: int, ctrlReg: int, ctrlQubit: int, theta) -> None:
        self._c("cvdvConditionalRotation", targetReg, ctrlReg, ctrlQubit, c_double(theta))

    def x(self, regIdx: int, targetQubit: int) -> None:
        self._c("cvdvPauliRotation", regIdx, targetQubit, 0, c_double(np.pi))

    def y(self, regIdx: int, targetQubit: int) -> None:
        self._c("cvdvPauliRotation", regIdx, targetQubit, 1, c_double(np.pi))

    def z(self, regIdx: int, targetQubit: int) -> None:
        self._c("cvdvPauliRotation", regIdx, targetQubit, 2, c_double(np.pi))

    def rx(self, regIdx: int, targetQubit: int, theta) -> None:
        self._c("cvdvPauliRotation", regIdx, targetQubit, 0, c_double(theta))

    def ry(self, regIdx: int, targetQubit: int, theta) -> None:
        self._c("cvdvPauliRotation", regIdx, targetQubit, 1, c_double(theta))

    def rz(self, regIdx: int, targetQubit: int, theta) -> None:
        self._c("cvdvPauliRotation", regIdx, targetQubit, 2, c_double(theta))

    def h(self, regIdx: int, targetQubit: int) -> None:
        self._c("cvdvHadamard", regIdx, targetQubit)

    def p(self, regIdx: int) -> None:
        self._c("cvdvParity", regIdx)

    def cp(self, targetReg: int, ctrlReg: int, ctrlQubit: int) -> None:
        self._c("cvdvConditionalParity", targetReg, ctrlReg, ctrlQubit)

    def swap(self, reg1: int, reg2: int) -> None:
        self._c("cvdvSwapRegisters", reg1, reg2)

    def sheer(self, regIdx: int, t) -> None:
        self._c("cvdvPhaseSquare", regIdx, c_double(t))

    def phaseCubic(self, regIdx: int, t) -> None:
        self._c("cvdvPhaseCubic", regIdx, c_double(t))

    def r(self, regIdx: int, theta) -> None:
        self._c("cvdvRotation", regIdx, c_double(theta))

    def s(self, regIdx: int, r) -> None:
        self._c("cvdvSqueeze", regIdx, c_double(r))

    def cs(self, targetReg: int, ctrlReg: int, ctrlQubit: int, r) -> None:
        self._c("cvdvConditionalSqueeze", targetReg, ctrlReg, ctrlQubit, c_double(r))

    def bs(self, reg1: int, reg2: int, theta) -> None:
        self._c("cvdvBeamSplitter", reg1, reg2, c_double(theta))

    def cbs(self, reg1: int, reg2: int, ctrlReg: int, ctrlQubit: int, theta) -> None:
        self._c("cvdvConditionalBeamSplitter", reg1, reg2, ctrlReg, ctrlQubit, c_double(theta))

    def q1q2(self, reg1: int, reg2: int, coeff) -> None:
        self._c("cvdvQ1Q2Gate", reg1, reg2, c_double(coeff))

    def ftQ2P(self, regIdx: int) -> None:
        self._c("cvdvFtQ2P", regIdx)

    def ftP2Q(self, regIdx: int) -> None:
        self._c("cvdvFtP2Q", regIdx)

    # ------------------------------------------------------------ readout
    def getState(self) -> npt.NDArray[np.complex128]:
        re = np.zeros(self.total_size, dtype=np.float64)
        im = np.zeros(self.total_size, dtype=np.float64)
        self._c("cvdvGetState", re.ctypes.data_as(POINTER(c_double)), im.ctypes.data_as(POINTER(c_double)))
        return re + 1j * im

    def getStateTensor(self) -> torch.Tensor:
        """The state as a complex128 device tensor (canonical layout); an addition."""
        t = torch.empty(self.total_size, dtype=torch.complex128, device="cuda")
        self._c("cvdvGetStateDevice", _dptr(t))
        return t

    def getXGrid(self, regIdx: int) -> npt.NDArray[np.float64]:
        dim = self.register_dims[regIdx]
        return (np.arange(dim) - (dim - 1) * 0.5) * self.grid_steps[regIdx]

    def m(self, *regIdxs) -> "float | npt.NDArray[np.float64]":
        """No argument: sum |psi|^2.  Registers: marginal array of shape (d_r0, d_r1, ...)."""
        if len(regIdxs) == 1 and isinstance(regIdxs[0], (list, tuple)):
            regIdxs = tuple(regIdxs[0])
        if not regIdxs:
            return float(self._c("cvdvGetNorm"))
        regs = list(regIdxs)
        out_size = int(np.prod([self.register_dims[r] for r in regs]))
        out = np.zeros(out_size, dtype=np.float64)
        regs_c = (c_int * len(regs))(*regs)
        self._c("cvdvMeasureMultiple", regs_c, c_int(len(regs)), out.ctypes.data_as(POINTER(c_double)))
        shape = tuple(self.register_dims[r] for r in regs)
        if len(shape) == 1:
            return out.reshape(shape)
        return out.reshape(shape[::-1]).T  # column-major C output, first register fastest

    def fidelityWith(self, other: "HipCvdv") -> float:
        out = c_double(0.0)
        self._lib.cvdvFidelityStatevectors(self.ctx, other.ctx, ctypes.byref(out))
        _lib.check()
        return float(out.value)

    def getFidelity(self, sep: SeparableState) -> float:
        arrays = self._device_arrays(sep)
        ptrs = (ctypes.c_void_p * self.num_registers)(*[_dptr(a) for a in arrays])
        out = c_double(0.0)
        self._c("cvdvGetFidelity", ptrs, c_int(self.num_registers), ctypes.byref(out))
        return float(out.value)

    def getPhotonNumber(self, regIdx: int) -> float:
        return float(self._c("cvdvGetPhotonNumber", c_int(regIdx)))

    # ------------------------------------------------------------ phase space (cudaCvdv.py:186-210)
    def _phase_axes(self, regIdx: int, p_factor: float):
        """Axes of the native N x N readouts: x_j = (j - (N-1)/2) dx and
        p_k = (k - N/2) p_factor pi / (N dx) (Wigner: p_factor 1; Husimi: 2)."""
        n = self.register_dims[regIdx]
        dx = float(self.grid_steps[regIdx])
        xs = dx * (np.arange(n) - 0.5 * (n - 1))
        ps = (p_factor * np.pi / (n * dx)) * (np.arange(n) - 0.5 * n)
        return xs, ps

    def _native_readout(self, entry: str, regIdx: int, bound, p_factor: float, device: bool):
        if not 0 <= regIdx < self.num_registers:
            raise ValueError(f"Invalid register index: {regIdx}")
        n = self.register_dims[regIdx]
        if device:
            out = torch.empty((n, n), dtype=torch.float64, device="cuda")
            self._c(entry + "Device", c_int(regIdx), _dptr(out))
        else:
            out = np.empty((n, n), dtype=np.float64)
            self._c(entry, c_int(regIdx), out.ctypes.data_as(POINTER(c_double)))
        if bound is None:
            return out
        xs, ps = self._phase_axes(regIdx, p_factor)
        rows, cols = np.flatnonzero(np.abs(ps) <= bound), np.flatnonzero(np.abs(xs) <= bound)
        if device:
            return out[torch.from_numpy(rows).to(out.device)][:, torch.from_numpy(cols).to(out.device)]
        return out[np.ix_(rows, cols)]

    def getWigner(self, regIdx: int, bound: float | None = None, device: bool = False):
        """Native-grid Wigner of register regIdx traced over the others: (N_p, N_x)
        rows momentum; `bound` keeps |x|, |p| <= bound.  device=True (an addition)
        returns a CUDA float64 tensor and skips the copy into host memory."""
        return self._native_readout("cvdvGetWigner", regIdx, bound, 1.0, device)

    def getHusimiQ(self, regIdx: int, bound: float | None = None, device: bool = False):
        """Native-grid Husimi-Q (route chosen as cvdvGetHusimiQ does); as getWigner."""
        return self._native_readout("cvdvGetHusimiQ", regIdx, bound, 2.0, device)

    def getHusimiQWigner(self, regIdx: int) -> npt.NDArray[np.float64]:
        """Husimi Q by Gaussian smoothing of the native Wigner grid (bosonic.cu:63-101)."""
        dim = self.register_dims[regIdx]
        out = np.zeros(dim * dim, dtype=np.float64)
        self._c("cvdvGetHusimiQWigner", regIdx, out.ctypes.data_as(POINTER(c_double)))
        return out.reshape((dim, dim))

    # ------------------------------------------------------------ additions
    def _grid(self, name, regIdx, a, b, per_slice, *extra):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        b = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1))
        if not 0 <= regIdx < self.num_registers:
            raise ValueError(f"Invalid register index: {regIdx}")
        S = self.total_size // self.register_dims[regIdx]
        blocks = S if per_slice else 1
        out = np.zeros(blocks * b.size * a.size, dtype=np.float64)
        self._c(name, c_int(regIdx), a.ctypes.data_as(POINTER(c_double)), c_int(a.size),
                b.ctypes.data_as(POINTER(c_double)), c_int(b.size), c_int(1 if per_slice else 0), *extra,
                out.ctypes.data_as(POINTER(c_double)))
        shape = (blocks, b.size, a.size) if per_slice else (b.size, a.size)
        return out.reshape(shape)

    def getWignerGrid(self, regIdx: int, xs, ps, per_slice: bool = False,
                      native: bool = True) -> npt.NDArray[np.float64]:
        """Wigner of register regIdx at the points (xs, ps): shape (len(ps), len(xs)), or
        (S, len(ps), len(xs)) per slice of the other registers (canonical order).
        native=True equals getWigner() on its grid (including its e^{-i(N-1)dx p}
        finalisation); native=False drops that phase (the symmetric sum)."""
        return self._grid("cvdvGetWignerGrid", regIdx, xs, ps, per_slice, c_int(1 if native else 0))

    def getHusimiQGrid(self, regIdx: int, qs, ps, per_slice: bool = False) -> npt.NDArray[np.float64]:
        """Husimi-Q of register regIdx at the points (qs, ps): shape (len(ps), len(qs)), or
        (S, len(ps), len(qs)) per slice; equals getHusimiQ() on its grid."""
        return self._grid("cvdvGetHusimiQGrid", regIdx, qs, ps, per_slice)

    def capture(self) -> "HipCvdvGraph":
        """Record the gate calls of a ``with`` block as a hipGraph instead of running
        them; ``launch()`` the result to apply them (replays compose)."""
        return HipCvdvGraph(self)

    def synchronize(self) -> None:
        self._c("cvdvSynchronize")

    def flush(self) -> None:
        """Apply deferred phases and restore the canonical register layout now."""
        self._c("cvdvFlush")

    def autotune(self) -> dict:
        """Time the FFT plans of every register (cvdvFftAutotune, an addition) and
        keep the fastest; returns {register: {"row_elems", "tile_cap", "fuse",
        "ms_per_transform"}} for the tuned registers (0 / -1: planner default)."""
        n = self._lib.cvdvFftAutotune(self.ctx)
        _lib.check()
        out = {"plans_timed": n}
        for r in range(self.num_registers):
            w, cap, fuse, ms = c_int(0), c_int(0), c_int(0), c_double(0.0)
            if self._lib.cvdvFftGetTuning(self.ctx, r, ctypes.byref(w), ctypes.byref(cap), ctypes.byref(fuse),
                                          ctypes.byref(ms)) == 1:
                out[r] = {"row_elems": w.value, "tile_cap": cap.value, "fuse": fuse.value,
                          "ms_per_transform": ms.value}
        return out

    def info(self) -> dict:
        return {"registers": self.qubit_counts.tolist(), "dims": self.register_dims,
                "grid_steps": self.grid_steps.tolist(), "total_size": self.total_size,
                "precision": self.precision}

    # ------------------------------------------------------------ plots (cudaCvdv.py:212-252)
    def _plot_field(self, field, xs, ps, bound, cmap, figsize, show, signed):
        import matplotlib.pyplot as plt

        rows, cols = np.abs(ps) <= bound, np.abs(xs) <= bound
        part = field[np.ix_(rows, cols)]
        span = {}
        if signed:  # a colour scale symmetric about zero
            top = float(np.max(np.abs(part)))
            span = {"vmin": -top, "vmax": top}
        fig, ax = plt.subplots(figsize=figsize)
        mesh = ax.pcolormesh(xs[cols], ps[rows], part, cmap=cmap, shading="auto", **span)
        ax.set_xlabel(r"$q$")
        ax.set_ylabel(r"$p$")
        fig.colorbar(mesh, ax=ax)
        fig.tight_layout()
        if show:
            plt.show()
        return fig, ax

    def plotWigner(self, regIdx: int, bound: float = 5.0, cmap: str = "RdBu",
                   figsize: Tuple[int, int] = (7, 6), show: bool = True) -> Tuple[Any, Any]:
        xs, ps = self._phase_axes(regIdx, 1.0)
        return self._plot_field(self.getWigner(regIdx), xs, ps, bound, cmap, figsize, show, signed=True)

    def plotHusimiQ(self, regIdx: int, bound: float = 5.0, cmap: str = "viridis",
                    figsize: Tuple[int, int] = (7, 6), show: bool = True) -> Tuple[Any, Any]:
        xs, ps = self._phase_axes(regIdx, 2.0)
        return self._plot_field(self.getHusimiQ(regIdx), xs, ps, bound, cmap, figsize, show, signed=False)


class HipCvdvGraph:
    """A gate sequence captured from a HipCvdv (see HipCvdv.capture)."""

    def __init__(self, sim: HipCvdv) -> None:
        self.sim = sim
        self.handle = None

    def __enter__(self) -> "HipCvdvGraph":
        self.sim._c("cvdvGraphBegin")
        return self

    def __exit__(self, exc_type, exc, tb) -> bool:
        handle = self.sim._lib.cvdvGraphEnd(self.sim.ctx)
        if exc_type is not None:
            if handle:
                self.sim._lib.cvdvGraphDestroy(handle)
            _lib.get_lib().cvdvGetLastError(None, 0)
            return False
        _lib.check()
        self.handle = handle
        return False

    def launch(self) -> None:
        if not self.handle:
            raise RuntimeError("no captured graph")
        self.sim._lib.cvdvGraphLaunch(self.handle)
        _lib.check()

    def __del__(self):
        try:
            if self.handle:
                self.sim._lib.cvdvGraphDestroy(self.handle)
                self.handle = None
        except Exception:
            pass
