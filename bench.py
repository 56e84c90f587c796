#!/usr/bin/env python3
"""Benchmark: Schwarz-screened FP64 Coulomb + exchange build on one MI355X.

Workload (default `c2_jk_fp64`): 40 atoms placed in a 14-bohr cube (seed 1002, every pair
>= 1.2 bohr apart), per atom 4 s (K=4), 3 p (K=3), 2 d (K=1), 1 f (K=1) spherical shells with
exponents log-uniform in [0.1, 1e3] and coefficients uniform in [0.1, 1] -> 1200 functions;
density D = C C^T with C 1200 x 100 of N(0, 0.1) entries; Schwarz threshold 1e-12.
A step is one J/K build (lbx_jk) with the basis, pair data, E matrices and Schwarz factors
already resident in HBM.  value = screened shell quartets per second.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "J/K build ms + ERI quartets/sec, mixed vs FP64 GPU at ≤1e-10 Eh; % FP64 peak"
FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
SCHWARZ = 1e-12


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(mol, D, gpu_stats, target_s: float):
    """The oracle (CPU FP64 restatement of the reference path, OpenMP over bra pairs) on a
    bounded sample: every stride-th bra pair of each class pair, stride chosen from a short
    calibration run so the sample takes about target_s seconds."""
    from oracle import oracle as O
    s = O.Structure(mol)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    stride = 20000
    _, _, st = s.jk(D, thresh=SCHWARZ, stride=stride)
    rate = st["quartets"] / max(st["seconds_jk"], 1e-9)
    want = target_s * rate
    stride = max(1, int(gpu_stats["quartets"] / max(want, 1.0)))
    _, _, st = s.jk(D, thresh=SCHWARZ, stride=stride)
    return {
        "value": st["quartets"] / st["seconds_jk"],
        "unit": "quartets/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"every {stride}th bra shell pair of each class pair: {st['quartets']:.0f} of "
                   f"{gpu_stats['quartets']:.0f} screened quartets, {st['seconds_jk']:.1f} s "
                   f"(+{st['seconds_setup']:.1f} s pair data and Schwarz factors, not counted)"),
        "gflops": st["flops"] / st["seconds_jk"] / 1e9,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2_jk_fp64")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from lible_amd import basis as B
    from lible_amd import lbx

    torch.cuda.set_device(0)
    cfg = B.CONFIGS[a.config]
    mol = B.config_molecule(a.config)
    t0 = time.perf_counter()
    ctx = lbx.Context(mol)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    Dh = B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"])
    D = torch.tensor(Dh, device="cuda:0")
    J = torch.empty_like(D)
    K = torch.empty_like(D)
    log(f"[bench] {a.config}: nbf={ctx.nbf} atoms={mol.natoms} setup {t_setup:.2f} s")

    t0 = time.perf_counter()
    _, _, st = ctx.jk(D, schwarz=SCHWARZ, J=J, K=K)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    if not a.no_tune:
        ctx.tune(D, schwarz=SCHWARZ)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    for _ in range(max(a.warmup, 1)):
        _, _, st = ctx.jk(D, schwarz=SCHWARZ, J=J, K=K)
    torch.cuda.synchronize()
    log(f"[bench] first call (Schwarz factors + task tables) {t_first:.2f} s, tuning {t_tune:.2f} s,"
        f" quartets {st['quartets']:.3e}")

    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        _, _, st = ctx.jk(D, schwarz=SCHWARZ, J=J, K=K)
        log(f"[bench] step {k} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)

    step_s = elapsed / a.steps
    value = st["quartets"] / step_s
    kern_ms = sum(p["ms"] for p in prof) / a.steps
    dom = max(prof, key=lambda p: p["ms"])
    dom_ms = dom["ms"] / dom["calls"]
    dom_tflops = dom["flops"] / dom["calls"] / (dom_ms * 1e-3) / 1e12
    tot_tflops = st["flops"] / step_s / 1e12
    # energy as a by-product check of the build
    E = 0.5 * float((D * J).sum()) - 0.25 * float((D * K).sum())

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": a.config,
            "atoms": mol.natoms, "box_bohr": cfg["box"], "nbf": ctx.nbf,
            "shells_per_atom": "4s(K4) 3p(K3) 2d(K1) 1f(K1) spherical",
            "density": f"C {ctx.nbf}x{cfg['nocc']} N(0,{cfg['sigma']})",
            "schwarz": SCHWARZ, "precision": "fp64",
            "quartets": st["quartets"], "flops_per_step": st["flops"],
        },
        "jk_build_ms": step_s * 1e3,
        "kernel_ms_per_step": kern_ms,
        "total_tflops": tot_tflops,
        "total_frac_fp64_peak": tot_tflops / FP64_PEAK_TFLOPS,
        "energy_2e": E,
        "roofline": {
            # FP64 compute bound: MI355X's 78.6 TFLOP/s FP64 peak is the same for MFMA and VALU
            "bound": "mfma",
            "kernel": "%s<%d,%d,%d,%d> J/K (%s)" % (
                "quartetM" if dom["kernel"] == "mfma" else "quartetV", *dom["class"],
                dom["kernel"]),
            "achieved": dom_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": dom_tflops / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel_ms": dom_ms,
            "kernel_share": dom["ms"] / max(sum(p["ms"] for p in prof), 1e-12),
        },
        "classes": sorted(
            [{"class": "".join(map(str, p["class"])), "kernel": p["kernel"],
              "ms": round(p["ms"] / p["calls"], 4),
              "quartets": p["quartets"] / p["calls"],
              "tflops": round(p["flops"] / max(p["ms"], 1e-9) / 1e9, 3)} for p in prof],
            key=lambda r: -r["ms"]),
        "setup_s": t_setup,
        "first_call_s": t_first,
        "tune_s": t_tune,
    }
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline (oracle sample)")
        out["cpu_baseline"] = cpu_baseline(mol, Dh, st, a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
