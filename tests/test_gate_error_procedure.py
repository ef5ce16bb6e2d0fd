"""The gate error studies of bench.py --gate-error (the reference's
analysis/{comm,disp,rot,squeeze,beam_splitter}_err.py procedure) run with the
FP64 oracle in place of the library: the complex128 scaling fits must come out
as the reference published them (ERROR-ANALYSIS.md:19-71), which pins the
analytic targets, the stop rule and the fits the GPU run uses.  The commutator
is compared in the variant with the reference's own Fock table (its
underflowing tails move the N = 1024 point)."""

import sys
import types

import numpy as np
import pytest

import bench
from oracle.cvdv_oracle import CvdvOracle

# ERROR-ANALYSIS.md: a(N) = c_a N^-1/2 + d_a, b(N) = c_b N^1/2 + d_b
PUBLISHED = {
    "commutator_reference_table": (6.3275, 0.1203, -7.8892, 30.4676),
    "displacement": (9.1432, -0.0733, -5.4037, 35.0233),
    "rotation": (5.0221, 0.1071, -5.6768, 18.2055),
    "squeeze": (5.9026, 0.4894, -1.7322, 7.8930),
    "beam_splitter": (4.9740, 0.0145, -4.3580, 6.9337),
}


class _OracleSim:
    def __init__(self, regs, precision="complex128"):
        self.o = CvdvOracle(regs, convention="torch")

    def initFromArray(self, a):
        self.o.initFromArray(np.asarray(a))

    def __getattr__(self, k):
        return getattr(self.o, k)


def test_published_fits_reproduced(monkeypatch):
    import torch

    fake = types.ModuleType("hipcvdv")
    fake.HipCvdv = _OracleSim
    monkeypatch.setitem(sys.modules, "hipcvdv", fake)
    monkeypatch.setattr(torch.cuda, "empty_cache", lambda: None)
    studies = {k: v for k, v in bench.GATE_STUDIES.items() if k in PUBLISHED}
    monkeypatch.setattr(bench, "GATE_STUDIES", studies)
    out = bench.gate_error_envelope(precs=("complex128",))
    for study, want in PUBLISHED.items():
        f = out[study]["fits"]["complex128"]
        got = (*f["a(N)"], *f["b(N)"])
        print(study, [round(v, 4) for v in got])
        assert got == pytest.approx(want, abs=0.01), study
