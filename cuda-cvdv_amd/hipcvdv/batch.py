"""HipCvdvBatch - many small independent circuits run by one kernel launch.

Each circuit of the batch is driven through a ``HipCvdv`` object whose gate
methods record instead of launching (same decompositions as a normal
simulator); ``run()`` then executes every circuit from its initial product
state in one launch, one workgroup per circuit.  Readouts: ``measure(reg)``
(per-circuit marginal of one register), ``norms()``, ``getState(i)``.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int
from typing import List, Sequence

import numpy as np
import numpy.typing as npt

from . import _lib
from .hipCvdv import PRECISIONS, HipCvdv


class HipCvdvBatch:
    """A batch of hybrid CV-DV circuits of (possibly) different register shapes."""

    def __init__(self, shapes: Sequence[Sequence[int]], precision: str = "complex128") -> None:
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = "complex128" if PRECISIONS[precision] == 0 else "mixed"
        self.shapes: List[List[int]] = [list(map(int, s)) for s in shapes]
        if not self.shapes:
            raise ValueError("a batch needs at least one circuit")
        self._lib = _lib.get_lib()
        nregs = (c_int * len(self.shapes))(*[len(s) for s in self.shapes])
        flat = [q for s in self.shapes for q in s]
        qflat = (c_int * max(1, len(flat)))(*flat)
        self.handle = self._lib.cvdvBatchCreate(len(self.shapes), nregs, qflat, PRECISIONS[precision])
        _lib.check()
        if not self.handle:
            raise _lib.CvdvError(2, "cvdvBatchCreate failed")
        self._circuits: List[HipCvdv] = []
        for i, s in enumerate(self.shapes):
            ctx = self._lib.cvdvBatchCircuit(self.handle, i)
            _lib.check()
            self._circuits.append(HipCvdv._recording(self._lib, ctx, s, self.precision, self))
        self._init_len = sum(1 << q for q in flat)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self._lib.cvdvBatchDestroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.shapes)

    def circuit(self, i: int) -> HipCvdv:
        """The recording simulator of circuit i (gate methods only)."""
        return self._circuits[i]

    def _c(self, name: str, *args):
        r = getattr(self._lib, name)(self.handle, *args)
        _lib.check()
        return r

    def initFromSeparable(self, states) -> None:
        """states[i]: a SeparableState or a list of per-register complex arrays."""
        if len(states) != len(self.shapes):
            raise ValueError(f"{len(states)} initial states for {len(self.shapes)} circuits")
        parts = []
        for i, (st, shape) in enumerate(zip(states, self.shapes)):
            regs = getattr(st, "register_arrays", st)
            if len(regs) != len(shape):
                raise ValueError(f"circuit {i}: {len(regs)} register arrays, expected {len(shape)}")
            for r, (a, q) in enumerate(zip(regs, shape)):
                if hasattr(a, "detach"):
                    a = a.detach().cpu().numpy()
                a = np.asarray(a, dtype=np.complex128).reshape(-1)
                if a.size != (1 << q):
                    raise ValueError(f"circuit {i} register {r}: {a.size} amplitudes, expected {1 << q}")
                parts.append(a)
        flat = np.ascontiguousarray(np.concatenate(parts)).view(np.float64)
        assert flat.size == 2 * self._init_len
        self._c("cvdvBatchInitFromSeparable", flat.ctypes.data_as(POINTER(c_double)))

    def setLaunch(self, threads: int, ldsBytes: int) -> None:
        """Threads per workgroup (256 or 512) and the LDS budget in bytes.  A state held
        in LDS takes 9/8 of its size (one pad slot per 8 amplitudes); larger states run
        from a global work area.  Results do not depend on the shape."""
        self._c("cvdvBatchSetLaunch", c_int(threads), c_int(ldsBytes))

    def getLaunch(self) -> tuple:
        t, l = c_int(0), c_int(0)
        self._c("cvdvBatchGetLaunch", ctypes.byref(t), ctypes.byref(l))
        return t.value, l.value

    def autotune(self, reps: int = 3) -> float:
        """Time every launch shape, keep the fastest; returns its microseconds per run."""
        return float(self._c("cvdvBatchAutotune", c_int(reps)))

    def numOps(self) -> int:
        return int(self._c("cvdvBatchNumOps"))

    def run(self) -> None:
        """Run every circuit from its initial state (asynchronous)."""
        self._c("cvdvBatchRun")

    def synchronize(self) -> None:
        self._c("cvdvBatchSynchronize")

    def measure(self, regIdx: int) -> List[npt.NDArray[np.float64]]:
        """Per circuit, the marginal probabilities of register regIdx."""
        sizes = [1 << s[regIdx] if regIdx < len(s) else 0 for s in self.shapes]
        out = np.zeros(max(1, sum(sizes)), dtype=np.float64)
        self._c("cvdvBatchMeasure", c_int(regIdx), out.ctypes.data_as(POINTER(c_double)))
        res, o = [], 0
        for n in sizes:
            res.append(out[o:o + n].copy())
            o += n
        return res

    def norms(self) -> npt.NDArray[np.float64]:
        out = np.zeros(len(self.shapes), dtype=np.float64)
        self._c("cvdvBatchNorms", out.ctypes.data_as(POINTER(c_double)))
        return out

    def getState(self, i: int) -> npt.NDArray[np.complex128]:
        n = int(np.prod([1 << q for q in self.shapes[i]]))
        re = np.zeros(n, dtype=np.float64)
        im = np.zeros(n, dtype=np.float64)
        self._c("cvdvBatchGetState", c_int(i), re.ctypes.data_as(POINTER(c_double)),
                im.ctypes.data_as(POINTER(c_double)))
        return re + 1j * im
