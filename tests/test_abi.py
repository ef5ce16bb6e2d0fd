"""CPU checks of the drop-in boundary: libcvdv_hip.so loads and exports exactly
what include/cvdv_hip.h declares, including every function of the reference's
src/cuda/api.h:14-61 (no compute calls: this container has no GPU)."""

import os
import re
import subprocess

import pytest

from hipcvdv import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cvdv_hip.h")

# src/cuda/api.h:14-61 of the reference (39 functions)
REFERENCE_API = """cvdvCreate cvdvDestroy cvdvFree cvdvInitFromSeparable cvdvSetStateFromDevicePtr
cvdvGetNumRegisters cvdvGetTotalSize cvdvGetRegisterInfo cvdvGetRegisterDim cvdvGetRegisterDx
cvdvFtQ2P cvdvFtP2Q cvdvDisplacement cvdvConditionalDisplacement cvdvPauliRotation cvdvHadamard
cvdvParity cvdvConditionalParity cvdvSwapRegisters cvdvPhaseSquare cvdvPhaseCubic cvdvRotation
cvdvConditionalRotation cvdvSqueeze cvdvConditionalSqueeze cvdvBeamSplitter
cvdvConditionalBeamSplitter cvdvQ1Q2Gate cvdvGetWigner cvdvGetHusimiQ cvdvGetHusimiQOverlap
cvdvGetHusimiQWigner cvdvMeasureMultiple cvdvMeasureMultipleCT cvdvGetState cvdvGetNorm
cvdvGetFidelity cvdvGetPhotonNumber cvdvFidelityStatevectors""".split()


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cvdv[A-Za-z0-9]+)\s*\(", txt)))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), "build libcvdv_hip.so first"


def test_header_covers_reference_api():
    assert len(REFERENCE_API) == 39
    missing = set(REFERENCE_API) - set(header_functions())
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    exp = exported_symbols()
    missing = [f for f in header_functions() if f not in exp]
    assert not missing, missing


def test_ctypes_signatures_match_header():
    assert set(_lib.SIGNATURES) == set(header_functions())
    lib = _lib.load()
    for name in header_functions():
        assert getattr(lib, name) is not None


def test_error_api_without_gpu():
    lib = _lib.get_lib()
    # A context with no registers is valid (context.cu:11-13) and touches no device.
    ctx = lib.cvdvCreate(0, None)
    assert ctx
    assert lib.cvdvGetNumRegisters(ctx) == 0
    assert lib.cvdvGetTotalSize(ctx) == 0
    assert lib.cvdvGetRegisterDim(ctx, 0) == -1
    lib.cvdvHadamard(ctx, 0, 0)  # uninitialised -> recorded error, no crash
    with pytest.raises(Exception):
        _lib.check()
    assert lib.cvdvGetLastError(None, 0) == 0  # cleared
    lib.cvdvDestroy(ctx)
