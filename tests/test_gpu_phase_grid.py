"""Wigner / Husimi-Q on arbitrary points (traced and per branch) against the oracle.

The oracle's grid readouts are pinned on the CPU (tests/test_oracle_closed_form.py:
equal to the native readouts on native points, closed forms off grid, per-slice
linearity); here the HIP path (rocBLAS GEMMs over FP64 operands) is compared with
them: complex128 to 1e-12 of the maximum, mixed to 2e-5 of the maximum (its
amplitudes carry relative errors of a few 1e-6 after 20 gates).
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(regs, prec, build):
    import torch
    from hipcvdv import HipCvdv, SeparableState
    from oracle.cvdv_oracle import CvdvOracle
    from oracle.separable_np import SeparableStateNP

    h = HipCvdv(regs, precision=prec)
    o = CvdvOracle(regs, convention="cuda")
    sep, sepn = SeparableState(regs), SeparableStateNP(regs)
    build(sep, sepn)
    h.initStateVector(sep)
    o.initStateVector(sepn)
    return h, o


def _qubit_oscillator_1_10(prec):
    """[1,10]: qubit |0>, coherent 2+1i; h, cd(1.5), r(0.7), s(0.3), h."""
    def build(sep, sepn):
        for s in (sep, sepn):
            s.setZero(0)
            s.setCoherent(1, 2 + 1j)
    h, o = _pair([1, 10], prec, build)
    for sim in (h, o):
        sim.h(0, 0)
        sim.cd(1, 0, 0, 1.5)
        sim.r(1, 0.7)
        sim.s(1, 0.3)
        sim.h(0, 0)
    return h, o


def _branches(prec, seed=5, nq=(2, 2), cv=9, ngates=20):
    """[2,2,cv]: qubits uniform, coherent oscillator, seeded cd / cr over the 4 qubits."""
    rng = np.random.default_rng(seed)
    alpha = complex(*rng.uniform(-1, 1, 2))

    def build(sep, sepn):
        for s in (sep, sepn):
            s.setUniform(0)
            s.setUniform(1)
            s.setCoherent(2, alpha)
    h, o = _pair([nq[0], nq[1], cv], prec, build)
    for _ in range(ngates):
        creg = int(rng.integers(0, 2))
        cq = int(rng.integers(0, nq[creg]))
        if rng.random() < 0.5:
            a = complex(*rng.uniform(-0.8, 0.8, 2))
            for sim in (h, o):
                sim.cd(2, creg, cq, a)
        else:
            th = float(rng.uniform(-math.pi, math.pi))
            for sim in (h, o):
                sim.cr(2, creg, cq, th)
    return h, o


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_wigner_201_grid_qubit_oscillator(prec):
    print(f"[grid] [1,10] Wigner on 201x201 {prec}", flush=True)
    h, o = _qubit_oscillator_1_10(prec)
    xs = ps = np.linspace(-5, 5, 201)
    tol = 1e-12 if prec == "complex128" else 2e-5
    for native in (True, False):
        Wg = h.getWignerGrid(1, xs, ps, native=native)
        Wo = o.getWignerGrid(1, xs, ps, native=native)
        assert Wg.shape == (201, 201)
        assert np.max(np.abs(Wg - Wo)) < tol * np.max(np.abs(Wo)), native
    Qg = h.getHusimiQGrid(1, xs, ps)
    Qo = o.getHusimiQGrid(1, xs, ps)
    assert np.max(np.abs(Qg - Qo)) < tol * np.max(Qo)
    np.testing.assert_allclose(h.m(0), o.m(0), atol=tol)


def test_grid_equals_native_readouts_on_native_points():
    print("[grid] native points", flush=True)
    h, _ = _branches("complex128", cv=8, ngates=8)
    N = 256
    dx = h.grid_steps[2]
    xv = (np.arange(N) - (N - 1) / 2) * dx
    pw = (np.arange(N) - N / 2) * math.pi / (N * dx)
    ph = (np.arange(N) - N / 2) * 2 * math.pi / (N * dx)
    ix, ip = np.arange(0, N, 5), np.arange(2, N, 7)
    W = h.getWigner(2)
    Q = h.getHusimiQ(2)
    np.testing.assert_allclose(h.getWignerGrid(2, xv[ix], pw[ip]), W[np.ix_(ip, ix)], atol=1e-13)
    np.testing.assert_allclose(h.getHusimiQGrid(2, xv[ix], ph[ip]), Q[np.ix_(ip, ix)], atol=1e-13)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_per_branch_against_oracle_and_linearity(prec):
    print(f"[grid] per-branch {prec}", flush=True)
    h, o = _branches(prec)
    xs = np.linspace(-6, 6, 37)
    ps = np.linspace(-6, 6, 41)
    tol = 1e-12 if prec == "complex128" else 2e-5
    Wb = h.getWignerGrid(2, xs, ps, per_slice=True)
    Qb = h.getHusimiQGrid(2, xs, ps, per_slice=True)
    assert Wb.shape == (16, 41, 37) and Qb.shape == (16, 41, 37)
    Wo = o.getWignerGrid(2, xs, ps, per_slice=True)
    Qo = o.getHusimiQGrid(2, xs, ps, per_slice=True)
    assert np.max(np.abs(Wb - Wo)) < tol * np.max(np.abs(Wo))
    assert np.max(np.abs(Qb - Qo)) < tol * np.max(Qo)
    np.testing.assert_allclose(Wb.sum(0), h.getWignerGrid(2, xs, ps), atol=1e-12)
    np.testing.assert_allclose(Qb.sum(0), h.getHusimiQGrid(2, xs, ps), atol=1e-12)
    assert np.all(Qb >= 0)


def test_512_grid_16_branches_subset_against_oracle():
    """[2,2,12], 16 branches, 512 x 512 over [-6,6]^2; the full grid on the
    GPU, checked against the oracle on 12 columns (every branch, every row)."""
    print("[grid] [2,2,12] 512x512 per branch", flush=True)
    h, o = _branches("complex128", seed=7, cv=12)
    xs = ps = np.linspace(-6, 6, 512)
    Wb = h.getWignerGrid(2, xs, ps, per_slice=True)
    Qb = h.getHusimiQGrid(2, xs, ps, per_slice=True)
    assert Wb.shape == (16, 512, 512)
    cols = np.arange(3, 512, 43)
    Wo = o.getWignerGrid(2, xs[cols], ps, per_slice=True)
    Qo = o.getHusimiQGrid(2, xs[cols], ps, per_slice=True)
    assert np.max(np.abs(Wb[:, :, cols] - Wo)) < 1e-12 * np.max(np.abs(Wo))
    assert np.max(np.abs(Qb[:, :, cols] - Qo)) < 1e-12 * np.max(Qo)
    probs = h.m(0, 1, 2)
    assert probs.shape == (4, 4, 4096)
    np.testing.assert_allclose(probs, o.m(0, 1, 2), atol=1e-13)


def test_grid_readout_errors():
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv([1, 6])
    sep = SeparableState([1, 6])
    sep.setZero(0)
    sep.setCoherent(1, 0.5)
    h.initStateVector(sep)
    with pytest.raises(ValueError):
        h.getWignerGrid(2, [0.0], [0.0])
    with pytest.raises(ValueError):
        h.getHusimiQGrid(1, [], [0.0])
    with pytest.raises(ValueError):
        h.getWignerGrid(1, [float("nan")], [0.0])
