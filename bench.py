"""Headline benchmark: gate passes/s and %HBM peak of a 2^28-amplitude hybrid circuit.

Workload (the reference's dense CV path at 2^28 amplitudes): two oscillator
registers of 2^14 grid points each ([14, 14]); reg 0 starts in the coherent
state alpha = 1.2 - 0.4i, reg 1 in squeezed vacuum r = 0.5; the circuit is 200
seeded Gaussian gates drawn uniformly from {displacement, rotation, squeeze,
50:50 beam splitter, random-angle beam splitter}.  One step = the whole
circuit applied to the state resident in HBM.

A gate pass is one algorithmic read+write of the state: every register
Fourier transform the gate decomposition needs (gates.cu:19-666) counts one;
diagonal phases are fused and count zero.  value = passes/s in the mixed mode
(complex64 amplitudes, FP64 angles/twiddles/sums); the FP64 (complex128) run,
the mixed/FP64 speedup and the fidelity loss of mixed against FP64 are
reported beside it.

The same JSON line carries the batched-circuit workload under "c4": 4096
independent circuits ([1-3 qubits, one oscillator of 2^8-2^12 points], 50-200
gates from {h, rx, rz, cd, r, s, cp}, seed 4) run by one launch, against the
same circuits issued one by one (per-gate kernel launches), and the mixed vs
complex128 agreement of their qubit marginals.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-c4]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cuda-cvdv_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

REGS = [14, 14]
N_GATES = 200
SEED = 2
HBM_PEAK_GBS = 8000.0
METRIC = "gate passes/s + %HBM peak at 2^28 amps; mixed-FP32 vs FP64 speedup @1e-6 err"


def make_circuit(seed=SEED, n=N_GATES):
    """200 Gaussian gates: (name, args).  |alpha| <= 1 (uniform disc), theta ~ U(-pi, pi),
    r ~ U(-0.3, 0.3), register uniform."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(n):
        kind = int(rng.integers(0, 5))
        reg = int(rng.integers(0, 2))
        if kind == 0:
            rad = math.sqrt(rng.random())
            ph = rng.uniform(0, 2 * math.pi)
            ops.append(("d", reg, complex(rad * math.cos(ph), rad * math.sin(ph))))
        elif kind == 1:
            ops.append(("r", reg, float(rng.uniform(-math.pi, math.pi))))
        elif kind == 2:
            ops.append(("s", reg, float(rng.uniform(-0.3, 0.3))))
        elif kind == 3:
            ops.append(("bs", 0, 1, math.pi / 2))
        else:
            ops.append(("bs", 0, 1, float(rng.uniform(-math.pi, math.pi))))
    return ops


def ft_count(op) -> int:
    """Register Fourier transforms issued by one gate (gates.cu:124-161, 316-348,
    394-421, 461-502, 572-636)."""
    name = op[0]
    if name == "d":
        return 2 if abs(op[2].real) > 1e-12 else 0
    if name == "s":
        return 4
    if name == "r":
        th = op[2]
        k = int(np.rint(th / (math.pi / 2)))
        rem = th - k * (math.pi / 2)
        return (1 if k % 4 in (1, 3) else 0) + (2 if abs(rem) >= 1e-15 else 0)
    if name == "bs":
        th = op[3]
        k = int(math.copysign(math.floor(abs(th / math.pi) + 0.5), th / math.pi))
        rem = th - k * math.pi
        return (2 if k % 4 in (1, 3) else 0) + (4 if abs(rem) >= 1e-15 else 0)
    raise ValueError(name)


def squeezed_vacuum(nq, r):
    N = 1 << nq
    dx = math.sqrt(2 * math.pi / N)
    x = (np.arange(N) - (N - 1) / 2) * dx
    return (math.exp(r / 2) * math.pi ** -0.25 * math.sqrt(dx) * np.exp(-0.5 * x * x * math.exp(2 * r))).astype(np.complex128)


# ---------------------------------------------------------------- batched circuits
C4_N = 4096
C4_SEED = 4


def coherent(nq, alpha):
    """pi^-1/4 sqrt(dx) e^{-(x-q)^2/2} e^{i(p x - p q/2)}, q = sqrt2 Re a, p = sqrt2 Im a."""
    N = 1 << nq
    dx = math.sqrt(2 * math.pi / N)
    x = (np.arange(N) - (N - 1) / 2) * dx
    q, p = math.sqrt(2) * alpha.real, math.sqrt(2) * alpha.imag
    return math.pi ** -0.25 * math.sqrt(dx) * np.exp(-0.5 * (x - q) ** 2) * np.exp(1j * (p * x - p * q / 2))


def make_c4(n=C4_N, seed=C4_SEED):
    """Circuits (shape, init, gates): shape [q, nb] with q ~ U{1,2,3} qubits as register 0 and
    an oscillator of 2^nb points, nb ~ U{8..12}, as register 1; the qubits start in |0..0>,
    the oscillator in a coherent state |alpha| <= 2 (uniform disc) or squeezed vacuum
    r ~ U(0, 0.8) (50/50); 50-200 gates uniform over {h, rx, rz, cd, r, s, cp} with
    angles ~ U(-pi, pi), cd |alpha| <= 0.5 (uniform disc), squeeze r ~ U(-0.3, 0.3)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        q = int(rng.integers(1, 4))
        nb = int(rng.integers(8, 13))
        if rng.random() < 0.5:
            rad, ph = 2 * math.sqrt(rng.random()), rng.uniform(0, 2 * math.pi)
            init = ("coherent", complex(rad * math.cos(ph), rad * math.sin(ph)))
        else:
            init = ("squeezed", float(rng.uniform(0, 0.8)))
        gates = []
        for _ in range(int(rng.integers(50, 201))):
            k = int(rng.integers(0, 7))
            qb = int(rng.integers(0, q))
            if k == 0:
                gates.append(("h", 0, qb))
            elif k == 1:
                gates.append(("rx", 0, qb, float(rng.uniform(-math.pi, math.pi))))
            elif k == 2:
                gates.append(("rz", 0, qb, float(rng.uniform(-math.pi, math.pi))))
            elif k == 3:
                rad, ph = 0.5 * math.sqrt(rng.random()), rng.uniform(0, 2 * math.pi)
                gates.append(("cd", 1, 0, qb, complex(rad * math.cos(ph), rad * math.sin(ph))))
            elif k == 4:
                gates.append(("r", 1, float(rng.uniform(-math.pi, math.pi))))
            elif k == 5:
                gates.append(("s", 1, float(rng.uniform(-0.3, 0.3))))
            else:
                gates.append(("cp", 1, 0, qb))
        out.append(([q, nb], init, gates))
    return out


def c4_init(shape, init):
    qa = np.zeros(1 << shape[0], dtype=np.complex128)
    qa[0] = 1.0
    cv = coherent(shape[1], init[1]) if init[0] == "coherent" else squeezed_vacuum(shape[1], init[1])
    return [qa, cv.astype(np.complex128)]


def c4_ft_count(g) -> int:
    if g[0] == "cd":
        return 2 if abs(g[4].real) > 1e-12 else 0
    if g[0] in ("r", "s"):
        return ft_count(g)
    return 0


def run_c4(prec, circuits, n_seq, steps):
    import torch
    from hipcvdv import HipCvdv, HipCvdvBatch

    n = len(circuits)
    batch = HipCvdvBatch([c[0] for c in circuits], precision=prec)
    t0 = time.perf_counter()
    for i, (_, _, gates) in enumerate(circuits):
        sim = batch.circuit(i)
        for g in gates:
            getattr(sim, g[0])(*g[1:])
    record_s = time.perf_counter() - t0
    batch.initFromSeparable([c4_init(c[0], c[1]) for c in circuits])
    tune_us = batch.autotune(reps=2)
    threads, lds = batch.getLaunch()

    def batched():
        batch.run()
        return batch.measure(0), batch.norms()

    batched()
    batch.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        probs, norms = batched()
    el = (time.perf_counter() - t0) / steps
    # the launch alone (every circuit, no readouts): best of 3
    kernel_s = 1e9
    for _ in range(3):
        t0 = time.perf_counter()
        batch.run()
        batch.synchronize()
        kernel_s = min(kernel_s, time.perf_counter() - t0)
    # the same circuits issued one by one: one context per shape, reused
    sims = {}
    seps = []
    for shape, init, _ in circuits[:n_seq]:
        if tuple(shape) not in sims:
            sims[tuple(shape)] = HipCvdv(shape, precision=prec)
        seps.append([torch.from_numpy(a).to("cuda") for a in c4_init(shape, init)])

    def seq_one(i):
        shape, _, gates = circuits[i]
        sim = sims[tuple(shape)]
        ptrs = [a.data_ptr() for a in seps[i]]
        import ctypes
        arr = (ctypes.c_void_p * 2)(*ptrs)
        sim._c("cvdvInitFromSeparable", arr, ctypes.c_int(2))
        for g in gates:
            getattr(sim, g[0])(*g[1:])
        return sim.m(0), sim.m()

    seq_one(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq_out = [seq_one(i) for i in range(n_seq)]
    seq_el = time.perf_counter() - t0
    agree = max(float(np.max(np.abs(seq_out[i][0] - probs[i]))) for i in range(n_seq))
    passes = sum(c4_ft_count(g) for c in circuits for g in c[2])
    amps = sum(1 << (c[0][0] + c[0][1]) for c in circuits)
    del sims, seps
    return {"batched_ms": el * 1e3, "kernel_ms": kernel_s * 1e3, "circuits_per_s": n / el, "record_s": record_s,
            "launch": {"threads": threads, "lds_bytes": lds, "autotuned_us": tune_us},
            "num_ops": batch.numOps(), "fft_passes": passes, "amplitudes": amps,
            "seq_circuits": n_seq, "seq_ms_per_circuit": seq_el / n_seq * 1e3,
            "seq_circuits_per_s": n_seq / seq_el,
            "speedup_vs_sequential": (seq_el / n_seq * n) / el,
            "seq_vs_batched_max_prob_diff": agree,
            "probs0": probs, "norms": norms}


# Sources each counter summary's kernels are compiled from (the phase programs
# they apply come from gates.hip).  A summary records their fingerprint; a figure
# is read from it only while the sources still match.
FFT_SOURCES = ["fft.hip", "fft_bfly.h", "diag_fast.h", "cvdv_internal.h", "gates.hip"]
BATCH_SOURCES = ["batch.hip", "fft_bfly.h", "diag_fast.h", "cvdv_internal.h", "gates.hip"]


def source_fingerprint(names):
    """sha256 (first 16 hex digits) of the named files of cuda-cvdv_amd/csrc, in order."""
    import hashlib
    h = hashlib.sha256()
    for n in names:
        with open(os.path.join(ROOT, "cuda-cvdv_amd", "csrc", n), "rb") as f:
            h.update(n.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def profile_figure(fname, sources, key):
    """(value, note): the figure `key` of the committed counter summary
    profiles/<fname>, or None with the reason (missing, or taken on other sources)."""
    try:
        with open(os.path.join(ROOT, "profiles", fname)) as f:
            lines = f.read().splitlines()
    except OSError:
        return None, f"profiles/{fname} missing"
    fp = next((l.split("=")[1].split()[0] for l in lines if l.startswith("sources_sha256 =")), None)
    if fp != source_fingerprint(sources):
        return None, f"profiles/{fname} was taken on other kernel sources (fingerprint {fp})"
    for l in lines:
        if l.startswith(key + " ="):
            return float(l.split("=")[1].split()[0]), None
    return None, f"{key} not in profiles/{fname}"


def c4_pmc(key):
    """A figure of the batch kernel from the committed counter summary of the C4
    workload (profiles/k_batch_c4_pmc.txt, separate rocprofv3 runs), or None."""
    return profile_figure("k_batch_c4_pmc.txt", BATCH_SOURCES, key)[0]


def c4_roofline(prec, r, circuits):
    """The batch kernel keeps each circuit's state on chip (LDS, or a work area in
    L2 / MALL): its algorithmic HBM bytes per launch are the inputs (one array per
    register, complex128) and the canonical outputs.  achieved = those bytes over
    the launch time; traffic = bytes per launch beyond the L2 (HBM and MALL) from
    counters, mostly the work areas' per-op reads and writes; valu_issue = VALU
    issue cycles over the SIMDs' cycles (counters).  Neither is near its peak: the
    kernel is latency-bound (two waves per SIMD, a barrier per Stockham stage)."""
    s = 8 if prec == "mixed" else 16
    out_b = sum((1 << (c[0][0] + c[0][1])) * s for c in circuits)
    in_b = sum(((1 << c[0][0]) + (1 << c[0][1])) * 16 for c in circuits)
    alg = out_b + in_b
    ach = alg / (r["kernel_ms"] * 1e-3) / 1e9
    tag = "c64" if prec == "mixed" else "c128"
    traffic, note = profile_figure("k_batch_c4_pmc.txt", BATCH_SOURCES, f"traffic_bytes_per_launch[{tag}]")
    return {"bound": "latency", "algorithmic_hbm_bytes": alg, "achieved_hbm_gbps": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": note,
            "traffic_gbps": traffic / (r["kernel_ms"] * 1e-3) / 1e9 if traffic else None,
            "valu_issue": c4_pmc(f"valu_issue_frac[{tag}]"), "kernel_ms": r["kernel_ms"]}


def calibration_copy():
    """A plain device copy of 2^26 complex64 amplitudes (512 MiB read and written):
    the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters in
    counter runs of a section that has no copy of its own."""
    import torch

    a = torch.zeros(1 << 26, dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    torch.cuda.empty_cache()


def c4_section(n, n_seq, steps):
    circuits = make_c4(n)
    calibration_copy()
    res = {}
    for prec in ("mixed", "complex128"):
        res[prec] = run_c4(prec, circuits, n_seq, steps)
        print(f"[bench] c4 {prec}: {res[prec]['batched_ms']:.2f} ms for {n} circuits, "
              f"{res[prec]['circuits_per_s']:.0f} circuits/s; sequential "
              f"{res[prec]['seq_ms_per_circuit']:.2f} ms/circuit; speedup "
              f"{res[prec]['speedup_vs_sequential']:.1f}x", file=sys.stderr, flush=True)
    pm, pd = res["mixed"].pop("probs0"), res["complex128"].pop("probs0")
    nm, nd = res["mixed"].pop("norms"), res["complex128"].pop("norms")
    for prec in ("mixed", "complex128"):
        res[prec]["roofline"] = c4_roofline(prec, res[prec], circuits)
    m = res["mixed"]
    return {
        "workload": f"{n} circuits [1-3 qubits, 2^8-2^12 oscillator], 50-200 gates "
                    "{h, rx, rz, cd, r, s, cp}, seed 4",
        "metric": "circuits/s, one batched launch vs sequential per-gate launches",
        "value": m["circuits_per_s"], "unit": "circuits/s", "precision": "mixed",
        "mixed": m, "complex128": res["complex128"],
        "mixed_vs_fp64_max_prob_diff": max(float(np.max(np.abs(a - b))) for a, b in zip(pm, pd)),
        "max_norm_dev_mixed": float(np.max(np.abs(nm - 1.0))),
        "max_norm_dev_fp64": float(np.max(np.abs(nd - 1.0))),
    }


# ---------------------------------------------------------------- state transfer (C3)
def make_c3(nq=14, ncv=14, seed=3):
    """Oscillator-to-qubit transfer: registers [1]*nq + [ncv] (oscillator last, its axis
    contiguous); the oscillator starts in a normalised superposition of 3 coherent
    states (|alpha| ~ U[0,3], arg ~ U[0,2pi), complex-normal coefficients), the qubits in
    |0>; round k = cd(osc, k, 0, beta_k), h(k, 0), cr(osc, k, 0, theta_k) with beta_k in
    the unit disc and theta_k ~ U(-pi, pi)."""
    rng = np.random.default_rng(seed)
    cat = []
    for _ in range(3):
        rad, ph = 3 * rng.random(), rng.uniform(0, 2 * math.pi)
        cat.append((complex(rad * math.cos(ph), rad * math.sin(ph)), complex(*rng.normal(size=2))))
    ops = []
    osc = nq
    for k in range(nq):
        rad, ph = math.sqrt(rng.random()), rng.uniform(0, 2 * math.pi)
        ops.append(("cd", osc, k, 0, complex(rad * math.cos(ph), rad * math.sin(ph))))
        ops.append(("h", k, 0))
        ops.append(("cr", osc, k, 0, float(rng.uniform(-math.pi, math.pi))))
    return [1] * nq + [ncv], cat, ops


def c3_section(steps):
    import torch
    from hipcvdv import HipCvdv, SeparableState

    regs, cat, ops = make_c3()
    nq = len(regs) - 1
    res = {}
    joint = {}
    for prec in ("mixed", "complex128"):
        h = HipCvdv(regs, precision=prec)
        sep = SeparableState(regs)
        for k in range(nq):
            sep.setZero(k)
        sep.setCat(nq, cat)

        def run():
            h.initStateVector(sep)
            for op in ops:
                getattr(h, op[0])(*op[1:])
            h.synchronize()

        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        t_circ = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        singles = [h.m(k) for k in range(nq)]
        t_single = time.perf_counter() - t0
        t0 = time.perf_counter()
        jp = h.m(*range(nq))
        t_joint = time.perf_counter() - t0
        ct = {}
        try:
            _ct_marginal(h, list(range(nq)))
            t0 = time.perf_counter()
            jct = _ct_marginal(h, list(range(nq)))
            ct = {"hiptensor_joint_ms": (time.perf_counter() - t0) * 1e3,
                  "hiptensor_joint_max_abs_diff": float(np.max(np.abs(jct - np.asarray(jp))))}
        except Exception as e:  # the library may not reduce a 15-mode tensor
            ct = {"hiptensor_joint": f"unsupported: {e}"}
        joint[prec] = (np.asarray(jp), [np.asarray(x) for x in singles])
        res[prec] = {"circuit_ms": t_circ * 1e3, "single_marginals_ms": t_single * 1e3,
                     "joint_marginal_ms": t_joint * 1e3, "joint_bins": int(np.asarray(jp).size),
                     "prob_sum": float(np.asarray(jp).sum()), **ct}
        print(f"[bench] c3 {prec}: circuit {t_circ * 1e3:.1f} ms, 14 marginals {t_single * 1e3:.1f} ms, "
              f"joint {t_joint * 1e3:.1f} ms", file=sys.stderr, flush=True)
        del h
        torch.cuda.empty_cache()
    jm, sm = joint["mixed"]
    jd, sd = joint["complex128"]
    return {"workload": "[1]x14 + [14] state transfer, 3-component cat, 14 rounds cd/h/cr, seed 3",
            "mixed": res["mixed"], "complex128": res["complex128"],
            "joint_max_abs_diff_mixed_vs_fp64": float(np.max(np.abs(jm - jd))),
            "single_max_abs_diff_mixed_vs_fp64": float(max(np.max(np.abs(a - b)) for a, b in zip(sm, sd))),
            "single_vs_joint_consistency_fp64": float(max(
                np.max(np.abs(sd[k] - jd.sum(axis=tuple(i for i in range(nq) if i != k)))) for k in range(nq)))}


# ---------------------------------------------------------------- phase-space readouts (C1, C5)
def make_c5(seed=5, cv=12, ngates=20):
    """[2,2,cv]: two 2-qubit registers in uniform superposition, an oscillator in a coherent
    state (seeded, |Re|,|Im| <= 1); ngates seeded gates, each a conditional displacement
    (|Re|,|Im| <= 0.8) or conditional rotation (theta ~ U(-pi, pi)) of the oscillator
    controlled by one of the 4 qubits."""
    rng = np.random.default_rng(seed)
    alpha = complex(*rng.uniform(-1, 1, 2))
    ops = []
    for _ in range(ngates):
        creg = int(rng.integers(0, 2))
        cq = int(rng.integers(0, 2))
        if rng.random() < 0.5:
            ops.append(("cd", 2, creg, cq, complex(*rng.uniform(-0.8, 0.8, 2))))
        else:
            ops.append(("cr", 2, creg, cq, float(rng.uniform(-math.pi, math.pi))))
    return [2, 2, cv], alpha, ops


def c5_section(steps):
    import torch
    from hipcvdv import HipCvdv, SeparableState

    regs, alpha, ops = make_c5()
    grid = np.linspace(-6, 6, 512)
    res, outs = {}, {}
    for prec in ("mixed", "complex128"):
        h = HipCvdv(regs, precision=prec)
        sep = SeparableState(regs)
        sep.setUniform(0)
        sep.setUniform(1)
        sep.setCoherent(2, alpha)
        h.initStateVector(sep)
        for op in ops:
            getattr(h, op[0])(*op[1:])
        h.synchronize()
        t = {}
        for name, fn in (("wigner_512_per_branch", lambda: h.getWignerGrid(2, grid, grid, per_slice=True)),
                         ("husimi_512_per_branch", lambda: h.getHusimiQGrid(2, grid, grid, per_slice=True)),
                         ("wigner_native_traced", lambda: h.getWigner(2)),
                         ("husimi_native_traced", lambda: h.getHusimiQ(2)),
                         # the same readouts into device memory (no 128 MiB copy to pageable host
                         # memory); the entries return once the result is complete
                         ("wigner_native_traced_device", lambda: h.getWigner(2, device=True)),
                         ("husimi_native_traced_device", lambda: h.getHusimiQ(2, device=True)),
                         ("joint_m012", lambda: h.m(0, 1, 2)),
                         ("joint_m012_hiptensor", lambda: _ct_marginal(h, [0, 1, 2]))):
            fn()  # untimed: first-call costs (code objects, library handles, work buffers)
            t0 = time.perf_counter()
            for _ in range(steps):
                val = fn()
            torch.cuda.synchronize()
            t[name + "_ms"] = (time.perf_counter() - t0) / steps * 1e3
            outs[(prec, name)] = val.cpu().numpy() if torch.is_tensor(val) else np.asarray(val)
        # device readouts: the time spent in the library's kernels, for the host-side share
        h._lib.cvdvProfileEnable(h.ctx, 1)
        h.getWigner(2, device=True)
        t["wigner_native_kernels_ms"] = _prof_total(h)
        h._lib.cvdvProfileReset(h.ctx)
        h.getHusimiQ(2, device=True)
        t["husimi_native_kernels_ms"] = _prof_total(h)
        h._lib.cvdvProfileEnable(h.ctx, 0)
        res[prec] = t
        print(f"[bench] c5 {prec}: " + ", ".join(f"{k} {v:.1f}" for k, v in t.items()), file=sys.stderr, flush=True)
        del h
        torch.cuda.empty_cache()
    err = {}
    for name in ("wigner_512_per_branch", "husimi_512_per_branch", "wigner_native_traced", "husimi_native_traced",
                 "joint_m012", "joint_m012_hiptensor"):
        a, b = outs[("mixed", name)], outs[("complex128", name)]
        err[name] = {"max_abs_err": float(np.max(np.abs(a - b))), "max_abs_fp64": float(np.max(np.abs(b)))}
    return {"workload": "[2,2,12], 16 qubit branches, 20 seeded cd/cr; per-branch Wigner and Husimi on "
                        "512x512 over [-6,6]^2, native traced readouts, joint m(0,1,2)",
            "mixed": res["mixed"], "complex128": res["complex128"], "mixed_vs_fp64": err}


def _ct_marginal(h, regs):
    """cvdvMeasureMultipleCT (the hipTensor comparator), shaped like HipCvdv.m()."""
    import ctypes

    from hipcvdv import _lib

    dims = [h.register_dims[r] for r in regs]
    out = np.zeros(int(np.prod(dims)))
    sel = (ctypes.c_int * len(regs))(*regs)
    h._lib.cvdvMeasureMultipleCT(h.ctx, sel, len(regs), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _lib.check()
    return out.reshape(dims[::-1]).T if len(dims) > 1 else out


def _prof_total(h):
    """Device time of the readouts recorded on the context: HIP events around each
    phase-space readout's kernels (profile kind 5), milliseconds."""
    from ctypes import byref, c_double

    tot, mx = c_double(0), c_double(0)
    h._lib.cvdvProfileRead(h.ctx, 5, byref(tot), byref(mx))
    return tot.value


def c1_section(steps):
    from hipcvdv import HipCvdv, SeparableState
    from oracle.cvdv_oracle import CvdvOracle
    from oracle.separable_np import SeparableStateNP

    grid = np.linspace(-5, 5, 201)

    def circuit(sim):
        sim.h(0, 0)
        sim.cd(1, 0, 0, 1.5)
        sim.r(1, 0.7)
        sim.s(1, 0.3)
        sim.h(0, 0)

    h = HipCvdv([1, 10])
    sep = SeparableState([1, 10])
    sep.setZero(0)
    sep.setCoherent(1, 2 + 1j)

    def gpu():
        h.initStateVector(sep)
        circuit(h)
        return h.m(0), h.getWignerGrid(1, grid, grid)

    gpu()
    t0 = time.perf_counter()
    for _ in range(steps):
        m0, W = gpu()
    t_gpu = (time.perf_counter() - t0) / steps
    o = CvdvOracle([1, 10], convention="cuda")
    sepn = SeparableStateNP([1, 10])
    sepn.setZero(0)
    sepn.setCoherent(1, 2 + 1j)
    t0 = time.perf_counter()
    o.initStateVector(sepn)
    circuit(o)
    mo, Wo = o.m(0), o.getWignerGrid(1, grid, grid)
    t_cpu = time.perf_counter() - t0
    return {"workload": "[1,10] complex128: h, cd(1.5), r(0.7), s(0.3), h; m(0) and the Wigner of the "
                        "oscillator on 201x201 over [-5,5]^2",
            "gpu_ms": t_gpu * 1e3, "cpu_oracle_ms": t_cpu * 1e3, "cpu_cores": 1,
            "max_abs_err_wigner": float(np.max(np.abs(W - Wo))), "max_abs_wigner": float(np.max(np.abs(Wo))),
            "max_abs_err_m0": float(np.max(np.abs(np.asarray(m0) - np.asarray(mo))))}


def run_gpu(prec, ops, steps, warmup, fuse=True, tune=True):
    """Time `steps` runs of the circuit on one context.  The timed region ends with
    cvdvFlush, which completes a half-done transform, applies the pending phases
    and restores the canonical layout: every gate of every step has executed.
    With `tune`, the FFT autotuner picks the tile plans first (outside the timed
    region; with fusion off it tunes tile shapes only)."""
    import ctypes

    import torch
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv(REGS, precision=prec)
    h._c("cvdvSetTransformFusion", ctypes.c_int(1 if fuse else 0))
    sep = SeparableState(REGS)
    sep.setCoherent(0, 1.2 - 0.4j)
    sep.setCoeffs(1, torch.from_numpy(squeezed_vacuum(REGS[1], 0.5)))
    h.initStateVector(sep)
    tuning = None
    if tune:
        t0 = time.perf_counter()
        tuning = h.autotune()
        if not fuse:  # the tuner may pick fusion for a field: keep every transform whole
            h._c("cvdvSetTransformFusion", ctypes.c_int(0))
        tuning["seconds"] = time.perf_counter() - t0
        print(f"[bench] {prec}: FFT autotune {tuning}", file=sys.stderr, flush=True)

    def step():
        for op in ops:
            getattr(h, op[0])(*op[1:])

    for _ in range(warmup):
        step()
    h.flush()
    h.synchronize()
    torch.cuda.synchronize()
    h._lib.cvdvProfileEnable(h.ctx, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    h.flush()
    h.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    from ctypes import byref, c_double
    tot, mx = c_double(0), c_double(0)
    prof = {}
    for kind, name in ((0, "fft"), (6, "fft_fused"), (7, "fft_row"), (1, "diag"), (2, "qubit"), (3, "cparity"),
                       (4, "reduce"), (5, "other")):
        n = h._lib.cvdvProfileRead(h.ctx, kind, byref(tot), byref(mx))
        prof[name + "_launches"] = n
        prof[name + "_ms"] = tot.value
    # the remaining FFT passes: four-step halves that do not fuse (first halves,
    # completions) and single-pass strided tiles
    prof["fft_half_launches"] = prof["fft_launches"] - prof["fft_fused_launches"] - prof["fft_row_launches"]
    prof["fft_half_ms"] = prof["fft_ms"] - prof["fft_fused_ms"] - prof["fft_row_ms"]
    h._lib.cvdvProfileEnable(h.ctx, 0)
    prof["tuning"] = tuning
    return h, el, prof


def fidelity_check(ops):
    """Fresh mixed and FP64 runs of one circuit: fidelity loss and max amplitude error."""
    import torch
    from hipcvdv import HipCvdv, SeparableState

    states = []
    sims = []
    for prec in ("complex128", "mixed"):
        h = HipCvdv(REGS, precision=prec)
        sep = SeparableState(REGS)
        sep.setCoherent(0, 1.2 - 0.4j)
        sep.setCoeffs(1, torch.from_numpy(squeezed_vacuum(REGS[1], 0.5)))
        h.initStateVector(sep)
        for op in ops:
            getattr(h, op[0])(*op[1:])
        sims.append(h)
    n0, n1 = sims[0].m(), sims[1].m()
    a = sims[0].getStateTensor()
    b = sims[1].getStateTensor().to(torch.complex128)
    # 1 - |<a|b>|^2 / (<a|a><b|b>), the three products summed alike (torch, FP64)
    ab, aa, bb = torch.vdot(a, b), torch.vdot(a, a).real, torch.vdot(b, b).real
    loss = float((1.0 - (ab.abs() ** 2) / (aa * bb)).item())
    err = float(torch.max(torch.abs(a - b)).item())
    scale = float(torch.max(torch.abs(a)).item())
    del a, b, sims
    torch.cuda.empty_cache()
    return loss, err / scale, n0, n1


def cpu_baseline(ops, budget_s):
    """The oracle (numpy/scipy FP64) on the same full 2^28 state, first gates of the
    same circuit until the time budget is used."""
    from oracle.cvdv_oracle import CvdvOracle
    from oracle.separable_np import SeparableStateNP

    workers = min(16, os.cpu_count() or 1)
    o = CvdvOracle(REGS, convention="cuda", workers=workers)
    sep = SeparableStateNP(REGS)
    sep.setCoherent(0, 1.2 - 0.4j)
    sep.setCoeffs(1, squeezed_vacuum(REGS[1], 0.5))
    o.initStateVector(sep)
    t0 = time.perf_counter()
    done, passes = 0, 0
    for op in ops:
        getattr(o, op[0])(*op[1:])
        done += 1
        passes += ft_count(op)
        if time.perf_counter() - t0 > budget_s and passes > 0:
            break
    el = time.perf_counter() - t0
    del o
    return {"value": passes / el, "unit": "passes/s", "cores": workers, "kind": "port",
            "sample": f"oracle (numpy/scipy FP64) on the full [14,14] state: first {done} gates "
                      f"({passes} passes) of the 200-gate circuit in {el:.1f} s, scipy.fft workers={workers}"}


def copy_probe(reps=10):
    """Plain device copy of one complex64 state (2^28 amplitudes, 2 GiB read + 2 GiB
    written per copy): the achievable-bandwidth reference for the FFT passes, and
    the known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters.
    Also the library's pattern-shaped copies (cvdvCopyProbe: each FFT pass's
    access pattern with no arithmetic), the per-pass ceilings."""
    import ctypes

    import torch
    from hipcvdv import _lib

    A = 1 << sum(REGS)
    src = torch.zeros(A, dtype=torch.complex64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    nbytes = 2 * A * 8
    patterns = {}
    lib = _lib.get_lib()
    for name, pat in _lib.PROBE_PATTERNS.items():
        v = ctypes.c_double(0.0)
        if lib.cvdvCopyProbe(pat, reps, ctypes.byref(v)) != 0:
            _lib.check()
        patterns[name] = {"ms": v.value, "gbps": nbytes / (v.value * 1e-3) / 1e9}
        print(f"[bench] copy probe {name}: {v.value:.3f} ms, {patterns[name]['gbps']:.0f} GB/s",
              file=sys.stderr, flush=True)
    return {"bytes_per_copy": nbytes, "ms": ms, "gbps": nbytes / (ms * 1e-3) / 1e9, "patterns": patterns}


def _bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = (u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return u.view(np.float32)


def dft_bf16_split_error(Ns=(64, 256, 1024), cols=64, seed=5):
    """Accuracy of the split-bf16 three-pass DFT-as-GEMM (matrix-core form): F_N
    and X each split as hi + lo bfloat16 (lo = bf16(v - hi)), Y = F_hi X_hi +
    F_hi X_lo + F_lo X_hi with the real and imaginary parts as real GEMMs, bf16
    products exact and summed in FP32 (the MFMA accumulator), against the FP64
    DFT.  Emulated on the CPU in numpy (the rounding is what the matrix cores
    would do; the FP32 accumulation order differs).  Returns, per N, the L2
    relative error over columns of unit-norm complex Gaussian data, to compare
    with the complex64 FFT's floor (the QFT envelope's 1.0-1.3e-7)."""
    rng = np.random.default_rng(seed)
    out = []
    for N in Ns:
        j = np.arange(N)
        F = np.exp(-2j * np.pi * np.outer(j, j) / N) / math.sqrt(N)
        X = rng.standard_normal((N, cols)) + 1j * rng.standard_normal((N, cols))
        X /= np.linalg.norm(X, axis=0)
        exact = F @ X

        def split(v):
            hi = _bf16(v.astype(np.float32))
            lo = _bf16((v - hi.astype(np.float64)).astype(np.float32))
            return hi.astype(np.float64), lo.astype(np.float64)

        def gemm3(A, B):  # real 3-pass product, products exact, sum rounded to FP32
            (ah, al), (bh, bl) = split(A), split(B)
            return (ah @ bh + ah @ bl + al @ bh).astype(np.float32).astype(np.float64)

        Fr, Fi, Xr, Xi = F.real, F.imag, X.real, X.imag
        Y = (gemm3(Fr, Xr) - gemm3(Fi, Xi)) + 1j * (gemm3(Fr, Xi) + gemm3(Fi, Xr))
        err = np.linalg.norm(Y - exact, axis=0) / np.linalg.norm(exact, axis=0)
        out.append({"N": N, "rel_l2_median": float(np.median(err)), "rel_l2_max": float(np.max(err))})
        print(f"[bench] split-bf16 DFT N={N}: relative L2 error median {np.median(err):.2e}, "
              f"max {np.max(err):.2e}", file=sys.stderr, flush=True)
    return out


def dft_gemm_probe(reps=5):
    """DFT-as-GEMM against the register FFT on one 2^28-amplitude complex64 state.

    For N in {64, 256, 1024}: the DFT of the leading axis of an (N, 2^28/N) view as
    one complex64 GEMM F_N @ X (rocBLAS/hipBLASLt through torch), the real bf16
    GEMM of the same shape that one term of a split-bf16 scheme needs (three such
    terms give FP32-like accuracy), and the library's FFT of an N-point register
    of a [28 - log2 N, log2 N] state (one pass, the register contiguous)."""
    import torch
    from hipcvdv import HipCvdv, SeparableState

    A = 1 << sum(REGS)
    out = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    for lg in (6, 8, 10):
        N = 1 << lg
        B = A // N
        j = torch.arange(N, device="cuda", dtype=torch.float64)
        F = torch.exp(-2j * math.pi * torch.outer(j, j) / N).to(torch.complex64) / math.sqrt(N)
        X = torch.randn(N, B, device="cuda", dtype=torch.complex64)
        Y = torch.empty_like(X)
        ms_c64 = timed(lambda: torch.matmul(F, X, out=Y))
        del X, Y
        torch.cuda.empty_cache()
        Fr = torch.randn(2 * N, 2 * N, device="cuda", dtype=torch.bfloat16)
        Xr = torch.randn(2 * N, B, device="cuda", dtype=torch.bfloat16)
        ms_bf16 = timed(lambda: torch.matmul(Fr, Xr))
        del Fr, Xr
        torch.cuda.empty_cache()
        regs = [sum(REGS) - lg, lg]
        h = HipCvdv(regs, precision="mixed")
        sep = SeparableState(regs)
        sep.setCoherent(0, 0.5)
        sep.setCoherent(1, 0.3)
        h.initStateVector(sep)
        h.ftQ2P(1)
        h.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            h.ftQ2P(1)
        h.synchronize()
        ms_fft = (time.perf_counter() - t0) / reps * 1e3  # the context's own stream
        del h
        torch.cuda.empty_cache()
        out[f"N={N}"] = {"gemm_c64_ms": ms_c64, "gemm_bf16_one_term_ms": ms_bf16,
                         "gemm_flops": 8.0 * N * N * B, "fft_pass_ms": ms_fft}
        print(f"[bench] dft-gemm N={N}: c64 GEMM {ms_c64:.2f} ms, bf16 GEMM (one of 3 split terms) "
              f"{ms_bf16:.2f} ms, register FFT {ms_fft:.3f} ms", file=sys.stderr, flush=True)
    return out


def _fock_table(N, shift=0.0, scale=1.0, pref=1.0):
    """Grid-sampled Hermite functions |n>, n < N (separable.py:62-74's recurrence
    for every n at once), FP64, at the coordinates x*scale - shift, times pref (the
    analytic displaced and squeezed Fock states of the gate studies).  The recurrence runs on mantissas m_n(x) with a
    per-point log scale l(x) (h_n = m_n e^l, l starting at -x^2/2), renormalised
    by 2^-500 when a mantissa passes 2^500, and each row is formed as
    sign(m) e^(log|m| + l): e^(-x^2/2) alone underflows for x^2/2 > 745, which
    would cut the tails of every |n> above n ~ 745 (grids of 2^11 points and up)."""
    dx = math.sqrt(2 * math.pi / N)
    x = (np.arange(N) - (N - 1) / 2) * dx * scale - shift
    F = np.empty((N, N))
    lscale = -0.5 * x * x
    cur = np.full(N, pref * math.pi ** -0.25 * math.sqrt(dx))
    prev = np.zeros(N)

    def row(m):
        with np.errstate(divide="ignore"):
            return np.sign(m) * np.exp(np.log(np.abs(m)) + lscale)

    F[0] = row(cur)
    for k in range(1, N):
        nxt = math.sqrt(2.0 / k) * x * cur - math.sqrt((k - 1.0) / k) * prev
        big = np.abs(nxt) > 2.0 ** 500
        if big.any():
            s = np.where(big, 2.0 ** -500, 1.0)
            nxt, cur = nxt * s, cur * s
            lscale = lscale + np.where(big, 500 * math.log(2.0), 0.0)
        prev, cur = cur, nxt
        F[k] = row(cur)
    return F


def _fock_table_plain(N):
    """The same Hermite functions by the plain recurrence from e^(-x^2/2), as the
    reference's analysis/_common.py:37-62 computes them: at N = 1024 the outermost
    points underflow to zero for every n."""
    dx = math.sqrt(2 * math.pi / N)
    x = (np.arange(N) - (N - 1) / 2) * dx
    F = np.empty((N, N))
    prev, cur = np.zeros(N), np.exp(-0.5 * x * x) * math.pi ** -0.25 * math.sqrt(dx)
    F[0] = cur
    for k in range(1, N):
        prev, cur = cur, math.sqrt(2.0 / k) * x * cur - math.sqrt((k - 1.0) / k) * prev
        F[k] = cur
    return F


def _upper_fit(ks, eps):
    """log eps = a k + b by least squares, b raised by the largest residual so the
    line bounds every point (the reference's fit_ab_upper, analysis/_common.py:67-86)."""
    if len(ks) < 2:
        return None
    y = np.log(np.asarray(eps))
    a, b = np.polyfit(np.asarray(ks, dtype=float), y, 1)
    b += float(np.max(y - (a * np.asarray(ks) + b)))
    return float(a), float(b)


def qft_error_envelope(qubits=range(5, 13)):
    """Precision envelope of the register QFT (the reference's analysis/qft_err.py
    and ERROR-ANALYSIS.md:13-16, rerun on this library): for N = 2^q, every Fock
    state |n> is transformed by ftQ2P (all n at once, as the slices of an [q, q]
    state) and compared with the analytic (-i)^n |n>; eps_n is the L2 error.  Per
    precision and N: the reference's upper-bound fit log eps <= a k + b over the
    points with eps >= 1e-9 (complex128) or >= 10x the mixed rounding floor, up to
    the first n with eps^2 >= 0.1; the floor (median eps over n < 8) and the
    largest error of mixed against complex128 on the same inputs.  Then
    a(N) = c_a N^-1/2 + d_a and b(N) = c_b N^1/2 + d_b, as in the reference."""
    import torch
    from hipcvdv import HipCvdv

    out = {"per_N": [], "fits": {}}
    fits = {"complex128": [], "mixed": []}
    for q in qubits:
        N = 1 << q
        F = _fock_table(N)
        target = F * ((-1j) ** (np.arange(N) % 4))[:, None]
        res = {}
        for prec in ("complex128", "mixed"):
            h = HipCvdv([q, q], precision=prec)
            h.initFromArray(F.astype(np.complex128).reshape(-1))
            h.ftQ2P(1)
            got = h.getState().reshape(N, N)
            del h
            torch.cuda.empty_cache()
            res[prec] = got
        row = {"N": N}
        for prec, got in res.items():
            eps = np.sqrt(np.sum(np.abs(got - target) ** 2, axis=1))
            stop = int(np.argmax(eps ** 2 >= 0.1)) if np.any(eps ** 2 >= 0.1) else N
            floor = float(np.median(eps[:8]))
            cut = 1e-9 if prec == "complex128" else 10 * floor
            ks = [k for k in range(stop + 1) if eps[k] >= cut]
            ab = _upper_fit(ks, eps[ks])
            row[prec] = {"floor": floor, "n_stop": stop, "fit_points": len(ks),
                         "a": ab[0] if ab else None, "b": ab[1] if ab else None}
            if ab:
                fits[prec].append((N, ab[0], ab[1]))
        d = np.sqrt(np.sum(np.abs(res["mixed"] - res["complex128"]) ** 2, axis=1))
        row["mixed_vs_c128_max_l2"] = float(np.max(d[: max(row["complex128"]["n_stop"], 1)]))
        out["per_N"].append(row)
        print(f"[bench] qft envelope N={N}: c128 floor {row['complex128']['floor']:.1e}, mixed floor "
              f"{row['mixed']['floor']:.1e}, mixed-vs-c128 {row['mixed_vs_c128_max_l2']:.1e}", file=sys.stderr,
              flush=True)
    # a(N) = c_a N^-1/2 + d_a, b(N) = c_b N^1/2 + d_b over every N swept, and over
    # the reference's own sweep (N = 32 .. 1024, analysis/qft_err.py:25), whose
    # published complex128 fit is a = 4.2125 N^-1/2 + 0.1027, b = -5.7903 N^1/2 + 16.2724
    for prec, fp in fits.items():
        for key, sel in (("all_N", fp), ("N_32_to_1024", [f for f in fp if 32 <= f[0] <= 1024])):
            if len(sel) >= 2:
                Ns = np.array([f[0] for f in sel], dtype=float)
                ca, da = np.polyfit(Ns ** -0.5, [f[1] for f in sel], 1)
                cb, db = np.polyfit(Ns ** 0.5, [f[2] for f in sel], 1)
                out["fits"].setdefault(prec, {})[key] = {"a(N)": [float(ca), float(da)],
                                                         "b(N)": [float(cb), float(db)]}
    return out


# The reference's other discretisation-error studies (ERROR-ANALYSIS.md:19-71,
# analysis/{comm,disp,rot,squeeze,beam_splitter}_err.py): grid sizes, gate
# parameters, the 0.1 stop and the 1e-9 fit cut-off as there.
GATE_STUDIES = {
    "commutator": {"N": [32, 64, 128, 256, 512, 1024]},
    # the same with the reference's underflowing Fock table (differs at N = 1024)
    "commutator_reference_table": {"N": [32, 64, 128, 256, 512, 1024], "plain": True},
    "displacement": {"N": [64, 128, 256, 512], "alpha": 4.0},
    "rotation": {"N": [64, 128, 256, 512], "theta": math.pi / 4},
    "squeeze": {"N": [64, 128, 256, 512], "r": 1.0, "min_n": 4},
    "beam_splitter": {"N": [32, 64, 128, 256], "theta": math.pi / 2, "min_n": 4},
}


def _gate_study_eps(study, N, prec):
    """Per-Fock errors eps_n of one study at grid size N: every Fock state at once
    as the slices of a state whose first register indexes n."""
    from hipcvdv import HipCvdv
    import torch

    q = N.bit_length() - 1
    par = GATE_STUDIES[study]
    F = _fock_table_plain(N) if par.get("plain") else _fock_table(N)

    def run(regs, arr, gate):
        h = HipCvdv(regs, precision=prec)
        h.initFromArray(arr.reshape(-1))
        gate(h)
        got = h.getState()
        del h
        torch.cuda.empty_cache()
        return got

    if study.startswith("commutator"):
        # ||x iQFT(x QFT psi) - iQFT(x QFT(x psi)) - i psi|| (comm_err.py:49-73)
        dx = math.sqrt(2 * math.pi / N)
        x = ((np.arange(N) - (N - 1) / 2) * dx)[None, :]
        qft = lambda a: run([q, q], a, lambda h: h.ftQ2P(1)).reshape(N, N)
        iqft = lambda a: run([q, q], a, lambda h: h.ftP2Q(1)).reshape(N, N)
        psi = F.astype(np.complex128)
        t1 = iqft(x * qft(psi))
        t2 = iqft(x * qft(x * psi))
        return np.linalg.norm(x * t1 - t2 - 1j * psi, axis=1)
    if study == "beam_splitter":
        # |n>|0> -> sum_k sqrt(C(n,k)) cos^k(t/2) (-i sin(t/2))^(n-k) |k>|n-k>
        # (beam_splitter_err.py:31-65)
        th = par["theta"]
        arr = F[:, :, None] * F[0][None, None, :]
        got = run([q, q, q], arr.astype(np.complex128), lambda h: h.bs(1, 2, th)).reshape(N, N, N)
        c, s_ = math.cos(th / 2), math.sin(th / 2)
        eps = np.empty(N)
        for n in range(N):
            k = np.arange(n + 1)
            lw = 0.5 * (math.lgamma(n + 1) - np.array([math.lgamma(i + 1) + math.lgamma(n - i + 1) for i in k]))
            with np.errstate(divide="ignore"):
                lw = lw + k * math.log(abs(c)) + (n - k) * math.log(abs(s_))
            w = np.exp(lw) * np.sign(c) ** k * ((-1j * np.sign(s_)) ** (n - k))
            target = (F[k].T * w) @ F[n - k]
            eps[n] = np.linalg.norm(got[n] - target)
        return eps
    if study == "displacement":
        a = par["alpha"]
        got = run([q, q], F.astype(np.complex128), lambda h: h.d(1, a)).reshape(N, N)
        target = _fock_table(N, shift=a * math.sqrt(2.0))
    elif study == "rotation":
        th = par["theta"]
        got = run([q, q], F.astype(np.complex128), lambda h: h.r(1, th)).reshape(N, N)
        target = F * np.exp(-1j * (np.arange(N) + 0.5) * th)[:, None]
    else:
        r = par["r"]
        got = run([q, q], F.astype(np.complex128), lambda h: h.s(1, r)).reshape(N, N)
        target = _fock_table(N, scale=math.exp(r), pref=math.exp(r / 2))
    return np.linalg.norm(got - target, axis=1)


def gate_error_envelope(precs=("complex128", "mixed")):
    """The reference's commutator, displacement D(4), rotation R(pi/4), squeeze
    S(1) and beam-splitter BS(pi/2) error studies rerun on this library in both
    precisions: eps_n per Fock state (the commutator's cumulative over n, as
    there), the upper-bound fit log eps <= a k + b over the points with eps >=
    1e-9 (complex128) or >= 10x the rounding floor (mixed) up to the first n with
    eps >= 0.1 (not before n = 4 where the reference waits), then a(N) = c_a N^-1/2
    + d_a and b(N) = c_b N^1/2 + d_b over the grid sizes."""
    out = {}
    for study, par in GATE_STUDIES.items():
        res = {"per_N": [], "fits": {}}
        fits = {p: [] for p in precs}
        for N in par["N"]:
            row = {"N": N}
            eps_by = {}
            for prec in precs:
                eps = _gate_study_eps(study, N, prec)
                if study.startswith("commutator"):
                    eps = np.cumsum(eps)
                eps_by[prec] = eps
                over = np.nonzero(eps >= 0.1)[0]
                over = over[over >= par.get("min_n", 0)]
                stop = int(over[0]) if len(over) else N - 1
                floor = float(np.median(eps[:8]))
                cut = 1e-9 if prec == "complex128" else 10 * floor
                ks = [k for k in range(stop + 1) if eps[k] >= cut]
                ab = _upper_fit(ks, eps[ks])
                row[prec] = {"floor": floor, "n_stop": stop, "fit_points": len(ks),
                             "a": ab[0] if ab else None, "b": ab[1] if ab else None}
                if ab:
                    fits[prec].append((N, ab[0], ab[1]))
            if len(precs) == 2:
                stop = row["complex128"]["n_stop"]
                row["mixed_vs_c128_max"] = float(np.max(np.abs(eps_by["mixed"][:stop + 1] -
                                                               eps_by["complex128"][:stop + 1])))
            res["per_N"].append(row)
            print(f"[bench] {study} N={N}: " + "; ".join(
                f"{p} stop n={row[p]['n_stop']} floor {row[p]['floor']:.1e}" for p in precs),
                file=sys.stderr, flush=True)
        for prec, fp in fits.items():
            if len(fp) >= 2:
                Ns = np.array([f[0] for f in fp], dtype=float)
                ca, da = np.polyfit(Ns ** -0.5, [f[1] for f in fp], 1)
                cb, db = np.polyfit(Ns ** 0.5, [f[2] for f in fp], 1)
                res["fits"][prec] = {"a(N)": [float(ca), float(da)], "b(N)": [float(cb), float(db)]}
        out[study] = res
    return out


def pmc_traffic(kind):
    """(HBM bytes per launch, note) of the FFT pass kind `kind` (fft_row, fft_half,
    fft_fused) from the committed counter summary of the [14,14] complex64
    Gaussian circuit (profiles/fft_c2_mixed_pmc.txt: FETCH_SIZE and WRITE_SIZE
    per kernel, each scaled by the same counter on a plain copy of known size);
    None and the reason when the summary is missing or was taken on other kernel
    sources.  The counters come from separate rocprofv3 runs, not from this one."""
    return profile_figure("fft_c2_mixed_pmc.txt", FFT_SOURCES, f"traffic_bytes_per_launch[{kind}]")


# FFT pass kinds of the profile (cvdvProfileRead) -> kernel name and the register
# transforms one launch performs (a fused pass completes one transform and starts
# the next: one transform's worth; a lone half: half of one; a one-row tile: one)
FFT_KINDS = {"fft_fused": ("k_fft_fused", 1.0), "fft_row": ("k_fft_tile (one-row)", 1.0),
             "fft_half": ("k_fft_tile (four-step half)", 0.5)}


def roofline(prof, pass_bytes, copy):
    """Roofline of the FFT pass kind with the largest summed time: algorithmic bytes
    per launch (transforms per launch x 2 A s) over its mean launch duration, both
    measured with HIP events on the context's stream in the timed region."""
    kinds = {k: v for k, v in FFT_KINDS.items() if prof.get(k + "_launches", 0) > 0}
    if not kinds:
        return None
    dom = max(kinds, key=lambda k: prof[k + "_ms"])
    name, per = FFT_KINDS[dom]
    n, ms = prof[dom + "_launches"], prof[dom + "_ms"]
    avg = ms / n
    alg = per * pass_bytes
    achieved = alg / (avg * 1e-3) / 1e9
    traffic, traffic_note = pmc_traffic(dom)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
           "kernel": name, "launches": n, "avg_launch_ms": avg, "algorithmic_bytes_per_launch": alg,
           "copy_gbps": copy["gbps"], "frac_of_copy": achieved / copy["gbps"], "share_of_fft_ms": ms / prof["fft_ms"],
           "fft_kinds": {k: {"kernel": FFT_KINDS[k][0], "launches": prof[k + "_launches"], "ms": prof[k + "_ms"],
                             "avg_ms": prof[k + "_ms"] / prof[k + "_launches"] if prof[k + "_launches"] else None,
                             "gbps": (FFT_KINDS[k][1] * pass_bytes / (prof[k + "_ms"] / prof[k + "_launches"] * 1e-3) / 1e9)
                             if prof[k + "_launches"] else None}
                         for k in FFT_KINDS}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only", choices=["mixed", "complex128"], default=None,
                    help="time a single precision (profiling runs)")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched-circuit workload")
    ap.add_argument("--c4-circuits", type=int, default=C4_N)
    ap.add_argument("--c4-seq", type=int, default=64, help="circuits timed one by one")
    ap.add_argument("--c4-only", action="store_true", help="run only the batched-circuit workload")
    ap.add_argument("--no-c3", action="store_true", help="skip the state-transfer workload")
    ap.add_argument("--c3-only", action="store_true", help="run only the state-transfer workload")
    ap.add_argument("--c15-only", action="store_true", help="run only the phase-space readout workloads (C1, C5)")
    ap.add_argument("--no-c5", action="store_true", help="skip the phase-space readout workloads (C1, C5)")
    ap.add_argument("--copy-only", action="store_true",
                    help="only the plain-copy bandwidth probe (counter calibration runs)")
    ap.add_argument("--qft-error", action="store_true",
                    help="only the QFT precision envelope (Fock states, both precisions)")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the precision-envelope fits (QFT and gate error studies)")
    ap.add_argument("--gate-error", action="store_true",
                    help="only the commutator, D, R, S and BS error studies (Fock states, both precisions)")
    ap.add_argument("--no-fuse", action="store_true",
                    help="run every register transform whole (two passes for the strided register)")
    ap.add_argument("--dft-bf16-error", action="store_true",
                    help="CPU emulation of the split-bf16 three-pass DFT-as-GEMM's accuracy")
    ap.add_argument("--dft-gemm", action="store_true",
                    help="only the DFT-as-GEMM (MFMA) against register-FFT comparison")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; the state vector is not sharded across GPUs.",
              file=sys.stderr)
        sys.exit(2)

    if args.dft_bf16_error:
        print(json.dumps({"dft_bf16_split_error": dft_bf16_split_error()}))
        return
    if args.dft_gemm:
        print(json.dumps({"dft_gemm": dft_gemm_probe()}))
        return
    if args.qft_error:
        print(json.dumps({"qft_error": qft_error_envelope()}))
        return
    if args.gate_error:
        print(json.dumps({"gate_error": gate_error_envelope()}))
        return
    if args.copy_only:
        print(json.dumps({"copy": copy_probe()}))
        return
    if args.c4_only:
        print(json.dumps(c4_section(args.c4_circuits, args.c4_seq, max(1, args.steps))))
        return
    if args.c3_only:
        print(json.dumps(c3_section(max(1, args.steps))))
        return
    if args.c15_only:
        print(json.dumps({"c1": c1_section(max(1, args.steps)), "c5": c5_section(max(1, args.steps))}))
        return
    ops = make_circuit()
    passes = sum(ft_count(op) for op in ops)
    A = 1 << sum(REGS)
    res = {}
    precs = [args.only] if args.only else ["mixed", "complex128"]
    for prec in precs:
        h, el, prof = run_gpu(prec, ops, args.steps, args.warmup, fuse=not args.no_fuse)
        s = 8 if prec == "mixed" else 16
        res[prec] = {"t": el, "prof": prof, "s": s}
        del h
        import torch
        torch.cuda.empty_cache()
        print(f"[bench] {prec}: {el / args.steps * 1e3:.1f} ms/step, {passes * args.steps / el:.1f} passes/s, "
              f"fft launches {prof['fft_launches']} in {prof['fft_ms']:.1f} ms", file=sys.stderr, flush=True)

    main_prec = precs[0]
    r = res[main_prec]
    value = passes * args.steps / r["t"]
    pass_bytes = 2 * A * r["s"]
    copy = copy_probe()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "passes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["t"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_prec == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "[14,14] oscillator registers, 2^28 amplitudes, 200 seeded Gaussian gates "
                               "(d, r, s, bs 50:50, bs random), seed 2",
                   "amplitudes": A, "passes_per_step": passes,
                   "fused_transforms": not args.no_fuse,
                   "fft_autotune": {str(k): v for k, v in (r["prof"].get("tuning") or {}).items()},
                   "fft_launches_per_step": r["prof"]["fft_launches"] / args.steps,
                   "precision": "complex64 amplitudes; FP64 angles, twiddles, sums" if main_prec == "mixed"
                   else "complex128"},
        "pct_hbm_peak": value * pass_bytes / 1e9 / HBM_PEAK_GBS * 100.0,
        "roofline": roofline(r["prof"], pass_bytes, copy),
        "copy_probe": copy,
    }
    if "complex128" in res and "mixed" in res:
        v64 = passes * args.steps / res["complex128"]["t"]
        out["fp64_value"] = v64
        out["fp64_pct_hbm_peak"] = v64 * 2 * A * 16 / 1e9 / HBM_PEAK_GBS * 100.0
        out["mixed_vs_fp64_speedup"] = value / v64
        loss, rel_err, n0, n1 = fidelity_check(ops)
        out["fidelity_loss_mixed_vs_fp64"] = loss
        out["max_rel_amp_err_mixed_vs_fp64"] = rel_err
        out["norm_fp64"] = n0
        out["norm_mixed"] = n1
    if not args.no_c3 and not args.only:
        out["c3"] = c3_section(max(1, args.steps))
    if not args.no_c5 and not args.only:
        out["c1"] = c1_section(max(1, args.steps))
        out["c5"] = c5_section(max(1, args.steps))
    if not args.no_c4 and not args.only:
        out["c4"] = c4_section(args.c4_circuits, args.c4_seq, max(1, args.steps))
    if not args.no_errors and not args.only:
        # the precision envelope (QFT and gate error studies): scaling fits only
        out["error_fits"] = {"qft": qft_error_envelope()["fits"],
                             "gates": {k: v["fits"] for k, v in gate_error_envelope().items()}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ops, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
