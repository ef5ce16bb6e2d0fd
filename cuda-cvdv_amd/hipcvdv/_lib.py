"""ctypes binding of libcvdv_hip.so (the C ABI declared in include/cvdv_hip.h).

The argument types mirror the reference binding in src/cudaCvdv.py:35-73, so
the same calls work on either library.  The library must have been built
(``python -c "import __graft_entry__ as g; g.build()"`` or ``make -C
cuda-cvdv_amd/csrc``); there is no fallback: a missing library raises.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_size_t, c_void_p
from typing import Optional

LIB_NAME = "libcvdv_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

_lib: Optional[ctypes.CDLL] = None

# name -> (argtypes, restype)
SIGNATURES = {
    "cvdvCreate": ([c_int, POINTER(c_int)], c_void_p),
    "cvdvCreateEx": ([c_int, POINTER(c_int), c_int], c_void_p),
    "cvdvDestroy": ([c_void_p], None),
    "cvdvFree": ([c_void_p], None),
    "cvdvInitFromSeparable": ([c_void_p, POINTER(c_void_p), c_int], None),
    "cvdvSetStateFromDevicePtr": ([c_void_p, c_void_p], None),
    "cvdvGetNumRegisters": ([c_void_p], c_int),
    "cvdvGetTotalSize": ([c_void_p], c_size_t),
    "cvdvGetRegisterInfo": ([c_void_p, POINTER(c_int), POINTER(c_double)], None),
    "cvdvGetRegisterDim": ([c_void_p, c_int], c_int),
    "cvdvGetRegisterDx": ([c_void_p, c_int], c_double),
    "cvdvFtQ2P": ([c_void_p, c_int], None),
    "cvdvFtP2Q": ([c_void_p, c_int], None),
    "cvdvDisplacement": ([c_void_p, c_int, c_double, c_double], None),
    "cvdvConditionalDisplacement": ([c_void_p, c_int, c_int, c_int, c_double, c_double], None),
    "cvdvPauliRotation": ([c_void_p, c_int, c_int, c_int, c_double], None),
    "cvdvHadamard": ([c_void_p, c_int, c_int], None),
    "cvdvParity": ([c_void_p, c_int], None),
    "cvdvConditionalParity": ([c_void_p, c_int, c_int, c_int], None),
    "cvdvSwapRegisters": ([c_void_p, c_int, c_int], None),
    "cvdvPhaseSquare": ([c_void_p, c_int, c_double], None),
    "cvdvPhaseCubic": ([c_void_p, c_int, c_double], None),
    "cvdvRotation": ([c_void_p, c_int, c_double], None),
    "cvdvConditionalRotation": ([c_void_p, c_int, c_int, c_int, c_double], None),
    "cvdvSqueeze": ([c_void_p, c_int, c_double], None),
    "cvdvConditionalSqueeze": ([c_void_p, c_int, c_int, c_int, c_double], None),
    "cvdvBeamSplitter": ([c_void_p, c_int, c_int, c_double], None),
    "cvdvConditionalBeamSplitter": ([c_void_p, c_int, c_int, c_int, c_int, c_double], None),
    "cvdvQ1Q2Gate": ([c_void_p, c_int, c_int, c_double], None),
    "cvdvGetWigner": ([c_void_p, c_int, POINTER(c_double)], None),
    "cvdvGetHusimiQ": ([c_void_p, c_int, POINTER(c_double)], None),
    "cvdvGetHusimiQOverlap": ([c_void_p, c_int, POINTER(c_double)], None),
    "cvdvGetHusimiQWigner": ([c_void_p, c_int, POINTER(c_double)], None),
    "cvdvGetWignerDevice": ([c_void_p, c_int, c_void_p], None),
    "cvdvGetHusimiQDevice": ([c_void_p, c_int, c_void_p], None),
    "cvdvSetFreeMemoryOverride": ([c_void_p, c_size_t], None),
    "cvdvMeasureMultiple": ([c_void_p, POINTER(c_int), c_int, POINTER(c_double)], None),
    "cvdvMeasureMultipleCT": ([c_void_p, POINTER(c_int), c_int, POINTER(c_double)], None),
    "cvdvMeasureMultipleNoAlloc": ([c_void_p, POINTER(c_int), c_int, POINTER(c_double)], None),
    "cvdvGetState": ([c_void_p, POINTER(c_double), POINTER(c_double)], None),
    "cvdvGetNorm": ([c_void_p], c_double),
    "cvdvGetFidelity": ([c_void_p, POINTER(c_void_p), c_int, POINTER(c_double)], None),
    "cvdvGetPhotonNumber": ([c_void_p, c_int], c_double),
    "cvdvFidelityStatevectors": ([c_void_p, c_void_p, POINTER(c_double)], None),
    "cvdvGetPrecision": ([c_void_p], c_int),
    "cvdvGetLastError": ([ctypes.c_char_p, c_size_t], c_int),
    "cvdvSynchronize": ([c_void_p], None),
    "cvdvFlush": ([c_void_p], None),
    "cvdvSetTransformFusion": ([c_void_p, c_int], None),
    "cvdvFftAutotune": ([c_void_p], c_int),
    "cvdvFftGetTuning": ([c_void_p, c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_double)],
                         c_int),
    "cvdvGetStateDevice": ([c_void_p, c_void_p], None),
    "cvdvProfileEnable": ([c_void_p, c_int], None),
    "cvdvProfileRead": ([c_void_p, c_int, POINTER(c_double), POINTER(c_double)], c_int),
    "cvdvProfileReset": ([c_void_p], None),
    "cvdvGraphBegin": ([c_void_p], None),
    "cvdvGraphEnd": ([c_void_p], c_void_p),
    "cvdvGraphLaunch": ([c_void_p], None),
    "cvdvGraphDestroy": ([c_void_p], None),
    "cvdvGetWignerGrid": ([c_void_p, c_int, POINTER(c_double), c_int, POINTER(c_double), c_int, c_int, c_int,
                           POINTER(c_double)], None),
    "cvdvGetHusimiQGrid": ([c_void_p, c_int, POINTER(c_double), c_int, POINTER(c_double), c_int, c_int,
                            POINTER(c_double)], None),
    "cvdvBatchCreate": ([c_int, POINTER(c_int), POINTER(c_int), c_int], c_void_p),
    "cvdvBatchDestroy": ([c_void_p], None),
    "cvdvBatchNumCircuits": ([c_void_p], c_int),
    "cvdvBatchCircuit": ([c_void_p, c_int], c_void_p),
    "cvdvBatchStateSize": ([c_void_p, c_int], c_size_t),
    "cvdvBatchNumOps": ([c_void_p], c_size_t),
    "cvdvBatchInitFromSeparable": ([c_void_p, POINTER(c_double)], None),
    "cvdvBatchSetLaunch": ([c_void_p, c_int, c_int], None),
    "cvdvBatchGetLaunch": ([c_void_p, POINTER(c_int), POINTER(c_int)], None),
    "cvdvBatchAutotune": ([c_void_p, c_int], c_double),
    "cvdvBatchRun": ([c_void_p], None),
    "cvdvBatchSynchronize": ([c_void_p], None),
    "cvdvBatchMeasure": ([c_void_p, c_int, POINTER(c_double)], None),
    "cvdvBatchNorms": ([c_void_p, POINTER(c_double)], None),
    "cvdvBatchGetState": ([c_void_p, c_int, POINTER(c_double), POINTER(c_double)], None),
    "cvdvCopyProbe": ([c_int, c_int, POINTER(c_double)], c_int),
}

# cvdvCopyProbe patterns (include/cvdv_hip.h)
PROBE_PATTERNS = {
    "float4 copy, grid 4 x CUs x 256": 0,
    "hipMemcpy device to device": 1,
    "strided FFT pass (128 x 256 B rows, stride 16 MiB, in place)": 2,
    "transposing FFT pass (rows stride 128 KiB -> stride 16 MiB)": 3,
    "mid-digit FFT pass (128 x 256 B rows, stride 128 KiB, in place)": 4,
    "one-row FFT pass (128 KiB row per workgroup, 1 per CU)": 5,
}


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the library and attach argtypes; raises if it has not been built."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"{LIB_NAME} not found at {path}; build it first "
            "(make -C cuda-cvdv_amd/csrc, or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (argt, rest) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = rest
    return lib


def get_lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


class CvdvError(RuntimeError):
    """A failure recorded by libcvdv_hip (see cvdvGetLastError)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libcvdv_hip error {code}: {msg}")
        self.code = code


ERR_ARG = 1


def check() -> None:
    """Raise the last recorded library error, if any (ValueError for bad arguments)."""
    lib = get_lib()
    buf = ctypes.create_string_buffer(512)
    code = lib.cvdvGetLastError(buf, 512)
    if code:
        msg = buf.value.decode(errors="replace")
        if code == ERR_ARG:
            raise ValueError(msg)
        raise CvdvError(code, msg)
