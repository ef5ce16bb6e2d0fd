"""Write golden vectors from the reference's own TorchCvdv (complex128, CPU).

Run once, in the build container only (the reference is not present on the GPU
box and never travels):

    python tests/golden/make_golden.py /root/reference

It imports ``src/torchCvdv.py`` and ``src/separable.py`` from the given reference
checkout with ``device='cpu'``, replays every case in ``CASES`` below and writes
``tests/golden/torchcvdv_golden.npz`` (inputs are the case descriptions
themselves, stored as JSON; outputs are the final flat states and readouts).
The fixtures pin the parts of the oracle that closed forms cannot: global
phases of the composite decompositions, the turn-count rounding at ties, the
Wigner (N-1) finalisation phase, the Husimi window at the grid edges, the m()
axis order and the init renormalisation.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PI = np.pi

# Each case: registers, separable init (list of [method, regIdx, *args]), ops
# (list of [method, *args]) and readouts (list of [method, *args]).
# Complex numbers are written as [re, im].
CASES = [
    # --- init builders
    {"name": "init_zero_uniform", "regs": [3, 4],
     "init": [["setZero", 0], ["setUniform", 1]], "ops": [], "reads": [["m", 0], ["m", 1, 0]]},
    {"name": "init_coherent", "regs": [6], "init": [["setCoherent", 0, [1.5, 0.8]]],
     "ops": [], "reads": [["m"]]},
    {"name": "init_fock3", "regs": [6], "init": [["setFock", 0, 3]], "ops": [], "reads": []},
    {"name": "init_focks", "regs": [6], "init": [["setFocks", 0, [[0.6, 0], [0, 0.8], [0, 0]]]],
     "ops": [], "reads": []},
    {"name": "init_cat", "regs": [7], "init": [["setCat", 0, [[[2.0, 0], [1.0, 0]], [[-2.0, 0], [1.0, 0]]]]],
     "ops": [], "reads": []},
    {"name": "init_coeffs_3reg", "regs": [1, 2, 3],
     "init": [["setCoeffs", 0, [[0.6, 0], [0, 0.8]]], ["setUniform", 1], ["setCoherent", 2, [0.3, -0.2]]],
     "ops": [], "reads": [["m", 2, 0]]},
    # --- Fourier transforms on both axes of a 2-register state
    {"name": "ft_q2p_reg0", "regs": [4, 3],
     "init": [["setCoherent", 0, [0.7, 0.4]], ["setCoherent", 1, [-0.3, 0.5]]],
     "ops": [["ftQ2P", 0]], "reads": []},
    {"name": "ft_q2p_reg1", "regs": [4, 3],
     "init": [["setCoherent", 0, [0.7, 0.4]], ["setCoherent", 1, [-0.3, 0.5]]],
     "ops": [["ftQ2P", 1]], "reads": []},
    {"name": "ft_p2q_reg0", "regs": [4, 3],
     "init": [["setCoherent", 0, [0.7, 0.4]], ["setCoherent", 1, [-0.3, 0.5]]],
     "ops": [["ftP2Q", 0]], "reads": []},
    {"name": "ft_roundtrip", "regs": [5],
     "init": [["setFock", 0, 2]], "ops": [["ftQ2P", 0], ["ftP2Q", 0]], "reads": []},
    # --- diagonal gates
    {"name": "sheer_cubic", "regs": [2, 5],
     "init": [["setUniform", 0], ["setCoherent", 1, [0.5, 0.2]]],
     "ops": [["sheer", 1, 0.13], ["phaseCubic", 1, -0.04], ["q1q2", 0, 1, 0.21]], "reads": []},
    {"name": "q1q2_rev", "regs": [3, 4],
     "init": [["setCoherent", 0, [0.5, 0]], ["setCoherent", 1, [0.8, 0]]],
     "ops": [["q1q2", 1, 0, -0.37]], "reads": []},
    # --- displacement family
    {"name": "d_complex", "regs": [6], "init": [["setZero", 0]], "ops": [["d", 0, [0.5, 0.3]]],
     "reads": []},
    {"name": "d_reg0_of_2", "regs": [5, 3],
     "init": [["setCoherent", 0, [0.2, 0.1]], ["setUniform", 1]],
     "ops": [["d", 0, [-0.6, 0.45]]], "reads": []},
    {"name": "cd_complex", "regs": [3, 4], "init": [["setUniform", 0], ["setZero", 1]],
     "ops": [["x", 0, 0], ["cd", 1, 0, 0, [0.5, 0.3]], ["cd", 1, 0, 2, [-0.2, 0.1]]],
     "reads": [["m", 0, 1]]},
    # --- rotation family (all four quarter-turn branches, plus a tie)
    {"name": "r_branches", "regs": [6], "init": [["setCoherent", 0, [1.0, 0.2]]],
     "ops": [["r", 0, 0.3], ["r", 0, 1.2], ["r", 0, 2.9], ["r", 0, -1.4]], "reads": []},
    {"name": "cr_branches", "regs": [3, 4], "init": [["setUniform", 0], ["setCoherent", 1, [0.8, 0]]],
     "ops": [["cr", 1, 0, 0, 0.4], ["cr", 1, 0, 1, 1.9], ["cr", 1, 0, 2, 3.3], ["cr", 1, 0, 0, -1.6]],
     "reads": []},
    {"name": "cr_tie_quarter", "regs": [1, 5], "init": [["setUniform", 0], ["setCoherent", 1, [0.8, 0]]],
     "ops": [["cr", 1, 0, 0, PI / 4]], "reads": []},
    # --- squeezing family
    {"name": "s_pos_neg", "regs": [6], "init": [["setZero", 0]],
     "ops": [["s", 0, 0.3], ["s", 0, -0.2]], "reads": []},
    {"name": "cs", "regs": [3, 4], "init": [["setZero", 0], ["setZero", 1]],
     "ops": [["x", 0, 0], ["h", 0, 1], ["cs", 1, 0, 0, 0.2], ["cs", 1, 0, 1, -0.15]], "reads": []},
    # --- beam splitter family (half-turn branches, plus a tie)
    {"name": "bs_branches", "regs": [3, 3],
     "init": [["setCoherent", 0, [0.5, 0]], ["setCoherent", 1, [0.8, 0.1]]],
     "ops": [["bs", 0, 1, 0.7], ["bs", 0, 1, 2.5], ["bs", 1, 0, -3.9], ["bs", 0, 1, 6.0]], "reads": []},
    {"name": "bs_tie_neg_half", "regs": [3, 3],
     "init": [["setCoherent", 0, [0.5, 0]], ["setCoherent", 1, [0.8, 0.1]]],
     "ops": [["bs", 0, 1, -PI / 2]], "reads": []},
    {"name": "cbs_branches", "regs": [2, 3, 3],
     "init": [["setUniform", 0], ["setCoherent", 1, [0.5, 0]], ["setCoherent", 2, [0.8, 0]]],
     "ops": [["cbs", 1, 2, 0, 0, 0.6], ["cbs", 1, 2, 0, 1, 2.8], ["cbs", 2, 1, 0, 0, -3.5]], "reads": []},
    # --- qubit gates and permutations
    {"name": "qubit_gates", "regs": [3, 2], "init": [["setZero", 0], ["setUniform", 1]],
     "ops": [["h", 0, 0], ["rx", 0, 1, 0.7], ["ry", 0, 2, -1.1], ["rz", 1, 0, 0.4],
             ["x", 0, 2], ["y", 1, 1], ["z", 0, 0], ["h", 1, 1]], "reads": [["m", 1, 0]]},
    {"name": "perms", "regs": [2, 3, 3],
     "init": [["setCoeffs", 0, [[0.1, 0.2], [0.3, -0.1], [0.5, 0.4], [-0.2, 0.6]]],
              ["setCoherent", 1, [0.4, 0.3]], ["setCoherent", 2, [-0.5, 0.2]]],
     "ops": [["p", 1], ["cp", 2, 0, 1], ["swap", 1, 2], ["cp", 1, 0, 0], ["p", 2]], "reads": []},
    {"name": "cp_same_register", "regs": [4], "init": [["setCoherent", 0, [0.4, 0.3]]],
     "ops": [["cp", 0, 0, 1]], "reads": []},
    # --- readouts
    {"name": "readouts_3reg", "regs": [2, 3, 4],
     "init": [["setUniform", 0], ["setCoherent", 1, [0.3, 0.2]], ["setFock", 2, 1]],
     "ops": [["cd", 2, 0, 0, [0.4, -0.3]], ["h", 0, 1], ["bs", 1, 2, 0.4]],
     "reads": [["m"], ["m", 0], ["m", 2], ["m", 2, 0], ["m", 1, 2, 0], ["m", 0, 1, 2],
               ["getPhotonNumber", 2], ["getWigner", 2], ["getHusimiQ", 2], ["getWigner", 1],
               ["getHusimiQ", 1]]},
    {"name": "wigner_husimi_cat", "regs": [5, 1],
     "init": [["setCat", 0, [[[1.0, 0.5], [1.0, 0]], [[-1.0, -0.5], [0, 1.0]]]], ["setUniform", 1]],
     "ops": [["cd", 0, 1, 0, [0.3, 0.2]]],
     "reads": [["getWigner", 0], ["getHusimiQ", 0], ["getPhotonNumber", 0]]},
    {"name": "fidelity_readout", "regs": [2, 4],
     "init": [["setUniform", 0], ["setFock", 1, 1]],
     "ops": [["d", 1, [0.3, 0.1]]],
     "reads": [["getFidelity", [["setUniform", 0], ["setFock", 1, 1]]]]},
]


def _rng_circuit(seed: int, regs, n_gates: int):
    """A seeded mixed circuit: qubit register 0, CV registers 1..."""
    rng = np.random.default_rng(seed)
    ops = []
    cv = list(range(1, len(regs)))
    nq0 = regs[0]
    for _ in range(n_gates):
        g = rng.integers(0, 9)
        t = int(rng.choice(cv))
        q = int(rng.integers(0, nq0))
        a = [float(rng.uniform(-0.6, 0.6)), float(rng.uniform(-0.6, 0.6))]
        th = float(rng.uniform(-3.0, 3.0))
        if g == 0:
            ops.append(["d", t, a])
        elif g == 1:
            ops.append(["cd", t, 0, q, a])
        elif g == 2:
            ops.append(["r", t, th])
        elif g == 3:
            ops.append(["cr", t, 0, q, th])
        elif g == 4:
            ops.append(["s", t, float(rng.uniform(-0.3, 0.3))])
        elif g == 5:
            ops.append(["h", 0, q])
        elif g == 6:
            ops.append(["rx", 0, q, th])
        elif g == 7 and len(cv) > 1:
            ops.append(["bs", cv[0], cv[1], th])
        else:
            ops.append(["cp", t, 0, q])
    return ops


CASES.append({"name": "random_circuit_2_4_4", "regs": [2, 4, 4],
              "init": [["setUniform", 0], ["setCoherent", 1, [0.4, -0.2]], ["setFock", 2, 1]],
              "ops": _rng_circuit(7, [2, 4, 4], 40), "reads": [["m", 0], ["m", 2, 1]]})


def _c(v):
    if isinstance(v, list) and len(v) == 2 and all(isinstance(u, (int, float)) for u in v):
        return complex(v[0], v[1])
    return v


def build_sep(SepCls, regs, init, **kw):
    sep = SepCls(regs, **kw)
    for item in init:
        meth, reg, *args = item
        if meth == "setCoherent":
            sep.setCoherent(reg, _c(args[0]))
        elif meth == "setFocks":
            sep.setFocks(reg, [_c(a) for a in args[0]])
        elif meth == "setCoeffs":
            sep.setCoeffs(reg, [_c(a) for a in args[0]])
        elif meth == "setCat":
            sep.setCat(reg, [(_c(a), _c(c)) for a, c in args[0]])
        else:
            getattr(sep, meth)(reg, *args)
    return sep


def apply_ops(sim, ops):
    for op in ops:
        meth, *args = op
        getattr(sim, meth)(*[_c(a) for a in args])


def run_reads(sim, reads, SepCls, regs, **kw):
    out = []
    for rd in reads:
        meth, *args = rd
        if meth == "m":
            out.append(np.asarray(sim.m(*args), dtype=np.float64))
        elif meth == "getFidelity":
            out.append(np.asarray(sim.getFidelity(build_sep(SepCls, regs, args[0], **kw))))
        else:
            out.append(np.asarray(getattr(sim, meth)(*args), dtype=np.float64))
    return out


def main(ref_root: str) -> None:
    sys.path.insert(0, ref_root)
    import torch  # noqa: F401  (the reference's backend)
    from src.separable import SeparableState
    from src.torchCvdv import TorchCvdv

    arrays = {}
    meta = []
    for case in CASES:
        sep = build_sep(SeparableState, case["regs"], case["init"], device="cpu")
        sim = TorchCvdv(case["regs"], device="cpu", dtype=np.complex128)
        sim.initStateVector(sep)
        for r, arr in enumerate(sep.register_arrays):
            arrays[f"{case['name']}/sep{r}"] = arr.numpy()
        arrays[f"{case['name']}/init"] = sim.getState()
        apply_ops(sim, case["ops"])
        arrays[f"{case['name']}/final"] = sim.getState()
        for i, v in enumerate(run_reads(sim, case["reads"], SeparableState, case["regs"], device="cpu")):
            arrays[f"{case['name']}/read{i}"] = v
        meta.append(case)
    arrays["cases_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    out = os.path.join(HERE, "torchcvdv_golden.npz")
    np.savez_compressed(out, **arrays)
    print(f"wrote {out}: {len(meta)} cases, {len(arrays)} arrays")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
