"""hipGraph capture of a gate sequence: every replay equals the same gates run
eagerly, with eager work and synchronisations between the replays."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sim(regs, prec):
    import torch
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv(regs, precision=prec)
    sep = SeparableState(regs)
    rng = np.random.default_rng(3)
    for r, q in enumerate(regs):
        a = rng.normal(size=1 << q) + 1j * rng.normal(size=1 << q)
        sep.setCoeffs(r, torch.from_numpy(a / np.linalg.norm(a)))
    h.initStateVector(sep)
    return h


def _gates(sim, regs):
    """A sequence touching every kind of pass: phases, qubit gate, FFTs (for [12, 2]
    the 12-qubit register is strided and needs the two-pass FFT, which ends in
    the other buffer), parity and swap in the lazy layout, a conditional parity."""
    q = len(regs) - 1 if regs[0] > regs[-1] else 0
    cv = 0 if q else len(regs) - 1
    sim.h(q, 0)
    sim.cd(cv, q, 0, 0.3 - 0.2j)
    sim.r(cv, 0.4)
    sim.rz(q, 0, 0.7)
    sim.s(cv, 0.15)
    sim.p(cv)
    sim.cp(cv, q, 0)
    sim.r(cv, 2.9)  # a quarter-turn branch
    if len(regs) == 3 and regs[1] == regs[2]:
        sim.bs(1, 2, 0.7)
        sim.swap(1, 2)


@pytest.mark.parametrize("regs", [[1, 8], [2, 6, 6], [12, 2]])
@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_graph_replays_match_eager(regs, prec):
    print(f"[graph] {regs} {prec}", flush=True)
    a = _sim(regs, prec)
    b = _sim(regs, prec)
    with a.capture() as g:
        _gates(a, regs)
    # nothing ran during the capture
    np.testing.assert_array_equal(a.getState(), b.getState())
    for rep in range(3):
        g.launch()
        _gates(b, regs)
        sa, sb = a.getState(), b.getState()  # readouts synchronise between replays
        scale = np.max(np.abs(sb))
        assert np.max(np.abs(sa - sb)) <= (1e-13 if prec == "complex128" else 1e-6) * scale, rep
        a.h(0, 0)  # eager work between replays
        b.h(0, 0)


def test_graph_capture_refuses_readouts_and_recovers():
    print("[graph] refusal", flush=True)
    a = _sim([1, 8], "complex128")
    before = a.getState()
    with pytest.raises(Exception):
        with a.capture():
            a.h(0, 0)
            a.m()  # readouts are refused while capturing
    # the failed capture left the state and the context usable
    np.testing.assert_array_equal(a.getState(), before)
    a.h(0, 0)
    with a.capture() as g:
        a.h(0, 0)
    g.launch()
    np.testing.assert_allclose(a.getState(), before, atol=1e-14)


def _many_controls(sim, nq):
    """More controlled phases on the oscillator than one kernel's tables hold (a
    spilled table pass), Rz on every qubit (shared one-bit tables), FFTs whose
    tables take the control bits as choice bits, and a qubit gate in between."""
    osc = nq
    for q in range(nq):
        sim.cd(osc, q, 0, 0.1j * (q + 1))
        sim.rz(q, 0, 0.3 + 0.2 * q)
    sim.ftQ2P(osc)
    sim.h(0, 0)
    for q in range(nq):
        sim.cd(osc, q, 0, 0.05 * (q + 1))
    sim.ftP2Q(osc)


@pytest.mark.parametrize("prec", ["complex128", "mixed"])
def test_graph_replays_shared_and_spilled_tables(prec):
    print(f"[graph] shared and spilled phase tables {prec}", flush=True)
    nq = 8
    regs = [1] * nq + [6]
    a = _sim(regs, prec)
    b = _sim(regs, prec)
    with a.capture() as g:
        _many_controls(a, nq)
    np.testing.assert_array_equal(a.getState(), b.getState())
    for rep in range(3):
        g.launch()
        _many_controls(b, nq)
        sa, sb = a.getState(), b.getState()
        scale = np.max(np.abs(sb))
        assert np.max(np.abs(sa - sb)) <= (1e-13 if prec == "complex128" else 1e-6) * scale, rep
        a.rz(1, 0, 0.4)  # eager work between replays
        b.rz(1, 0, 0.4)
