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

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
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


def run_gpu(prec, ops, steps, warmup):
    import torch
    from hipcvdv import HipCvdv, SeparableState

    h = HipCvdv(REGS, precision=prec)
    sep = SeparableState(REGS)
    sep.setCoherent(0, 1.2 - 0.4j)
    sep.setCoeffs(1, torch.from_numpy(squeezed_vacuum(REGS[1], 0.5)))
    h.initStateVector(sep)

    def step():
        for op in ops:
            getattr(h, op[0])(*op[1:])

    for _ in range(warmup):
        step()
    h.synchronize()
    torch.cuda.synchronize()
    h._lib.cvdvProfileEnable(h.ctx, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    h.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    from ctypes import byref, c_double
    tot, mx = c_double(0), c_double(0)
    n_fft = h._lib.cvdvProfileRead(h.ctx, 0, byref(tot), byref(mx))
    prof = {"fft_launches": n_fft, "fft_ms": tot.value}
    for kind, name in ((1, "diag"), (2, "qubit"), (4, "reduce"), (5, "other")):
        n = h._lib.cvdvProfileRead(h.ctx, kind, byref(tot), byref(mx))
        prof[name + "_launches"] = n
        prof[name + "_ms"] = tot.value
    h._lib.cvdvProfileEnable(h.ctx, 0)
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
    f = sims[1].fidelityWith(sims[0])
    n0, n1 = sims[0].m(), sims[1].m()
    loss = 1.0 - f / (n0 * n1)
    a = sims[0].getStateTensor()
    b = sims[1].getStateTensor()
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only", choices=["mixed", "complex128"], default=None,
                    help="time a single precision (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; the state is not sharded across GPUs "
              "(see DESIGN.md, 'Sharding').", file=sys.stderr)
        sys.exit(2)

    ops = make_circuit()
    passes = sum(ft_count(op) for op in ops)
    A = 1 << sum(REGS)
    res = {}
    precs = [args.only] if args.only else ["mixed", "complex128"]
    for prec in precs:
        h, el, prof = run_gpu(prec, ops, args.steps, args.warmup)
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
    fft_alg_bytes = passes * args.steps * pass_bytes
    fft_launches = r["prof"]["fft_launches"]
    fft_ms = r["prof"]["fft_ms"]
    achieved = fft_alg_bytes / (fft_ms * 1e-3) / 1e9 if fft_ms > 0 else None
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
                   "precision": "complex64 amplitudes; FP64 angles, twiddles, sums" if main_prec == "mixed"
                   else "complex128"},
        "pct_hbm_peak": value * pass_bytes / 1e9 / HBM_PEAK_GBS * 100.0,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "kernel": "k_fft_tile", "fft_launches": fft_launches,
                     "avg_launch_ms": fft_ms / fft_launches if fft_launches else None,
                     "algorithmic_bytes_per_launch": fft_alg_bytes / fft_launches if fft_launches else None},
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
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ops, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
