#!/usr/bin/env python3
"""Benchmark: Schwarz-screened FP64 Coulomb + exchange build on one MI355X.

Workload (default `c2_jk_fp64`): 40 atoms placed in a 14-bohr cube (seed 1002, every pair
>= 1.2 bohr apart), per atom 4 s (K=4), 3 p (K=3), 2 d (K=1), 1 f (K=1) spherical shells with
exponents log-uniform in [0.1, 1e3] and coefficients uniform in [0.1, 1] -> 1200 functions;
density D = C C^T with C 1200 x 100 of N(0, 0.1) entries; Schwarz threshold 1e-12.
A step is one J/K build (lbx_jk) with the basis, pair data, E matrices and Schwarz factors
already resident in HBM.  value = screened shell quartets per second of the FP64 build.
After the FP64 steps, --mixed-steps builds of the mixed-precision form (FP32 R and transforms
for quartets with Schwarz bound < 1e-5) are timed and compared with the FP64 result ("mixed").
--config c3_jk_mixed runs the 120-atom, 3600-function workload instead; --config c4_ri_batch
the batched RI integrals of 256 molecules (value = (mn|P) triplets/s); --config c5_gradient
the nuclear gradient of the 60-atom workload (value = quartets/s of the gradient).

Prints one JSON line.  --gpus N > 1 exits: the benchmark runs on one GPU.
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
MIXED_SPLIT = 1e-5           # mixed precision: FP32 for quartets with Q_ab Q_cd below this


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
    for _ in range(3):
        # small samples run below the steady rate (thread imbalance): correct from the last one
        if st["seconds_jk"] >= 0.5 * target_s or stride == 1:
            break
        rate = st["quartets"] / max(st["seconds_jk"], 1e-9)
        stride = max(1, int(gpu_stats["quartets"] / max(target_s * rate, 1.0)))
        _, _, st = s.jk(D, thresh=SCHWARZ, stride=stride)
    return {
        "value": st["quartets"] / st["seconds_jk"],
        "unit": "quartets/s",
        "cores": cores,
        "kind": "port",
        "path": "the oracle's restatement of the reference's generic-L CPU path "
                "(eri_kernel_funs.cpp:7-101, triple-loop GEMMs; not its rolled-out L <= 6 kernels)",
        "sample": (f"every {stride}th bra shell pair of each class pair: {st['quartets']:.0f} of "
                   f"{gpu_stats['quartets']:.0f} screened quartets, {st['seconds_jk']:.1f} s "
                   f"(+{st['seconds_setup']:.1f} s pair data and Schwarz factors, not counted)"),
        "gflops": st["flops"] / st["seconds_jk"] / 1e9,
    }


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs
    (profiles/jk_c2_traffic.json: FETCH_SIZE and WRITE_SIZE passes, each scaled by the same
    counter's reading for a copy of known size), or None when they cover another kernel."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "jk_c2_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes") if t.get("kernel") == kernel else None


def energy_difference(D, J, K, Jm, Km) -> float:
    """E(Jm, Km) - E(J, K) for E = 1/2 tr(DJ) - 1/4 tr(DK), formed from the differences
    1/2 tr(D (Jm - J)) - 1/4 tr(D (Km - K)) so that it is not lost in the rounding of two
    energies of ~1e5 Eh (one ulp of such an energy is ~1e-11 Eh)."""
    return 0.5 * float((D * (Jm - J)).sum()) - 0.25 * float((D * (Km - K)).sum())


def kernel_name(prof: dict) -> str:
    """The template name of a class pair's J/K kernel from its profile record."""
    kind = prof["kernel"]
    base = "quartetM" if kind.startswith("mfma") else "quartetB" if kind.startswith("block") else "quartetV"
    return "%s<%d,%d,%d,%d> J/K (%s)" % (base, *prof["class"], kind)


def timed(fn, steps: int, label: str):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = None
    for k in range(steps):
        res = fn()
        log(f"[bench] {label} step {k} done")
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, res


CLASS_PAIRS = [(la, lb, lc, ld) for la in range(4) for lb in range(la + 1)
               for lc in range(4) for ld in range(lc + 1)
               if la * (la + 1) // 2 + lb >= lc * (lc + 1) // 2 + ld]


def fp32_tier_error(ctx, D, per_class: bool, split: float = MIXED_SPLIT) -> dict:
    """The error of the FP32 tier alone: the quartets with Schwarz bound in [SCHWARZ, split)
    built in FP64 and in FP32 (lbx_set_bound_window), so the partial J and K are small and their
    FP64 rounding does not hide the difference; per class pair with per_class (max |dJ|, max |dK|
    and the energy change of each class pair's FP32 quartets)."""
    from lible_amd import lbx

    def one():
        J, K, st = ctx.jk(D, schwarz=SCHWARZ)
        Jm, Km, sm = ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED, mixed_thresh=split)
        r = {"split": split, "quartets": sm["quartets_fp32"], "delta_E": energy_difference(D, J, K, Jm, Km),
             "max_abs_dJ": float((Jm - J).abs().max()), "max_abs_dK": float((Km - K).abs().max())}
        return r

    ctx.set_bound_window(0.0, split)
    try:
        res = one()
        if per_class:
            rows = []
            for c in CLASS_PAIRS:
                ctx.set_class_filter(c)
                r = one()
                if r["quartets"] > 0:
                    rows.append({"class": "".join(map(str, c)), **r})
            ctx.set_class_filter(None)
            res["per_class"] = rows
    finally:
        ctx.set_class_filter(None)
        ctx.set_bound_window()
    return res


def leg_c1() -> dict:
    """Dense eri4 of the 6-atom, 84-function workload (3 s(K3), 2 p(K2), 1 d per atom)."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    mol = B.config_molecule("c1_eri4_dense")
    ctx = lbx.Context(mol)
    n = ctx.nbf
    out = torch.empty((n, n, n, n), dtype=torch.float64, device="cuda")
    ctx.eri4(out)
    torch.cuda.synchronize()
    t, _ = timed(lambda: ctx.eri4(out), 3, "c1 eri4")
    s = float(out.abs().sum())
    ctx.close()
    log("[bench] c1 done")
    return {"workload": "c1_eri4_dense", "nbf": n, "ms": t * 1e3, "elements_per_s": n ** 4 / t,
            "sum_abs": s}


C3_SPLIT = 1e-6   # at 120 atoms the FP32 rounding of the quartets below 1e-5 adds up to ~3e-10 Eh
TUNED = {}        # kernel per class pair chosen by the autotuner on the main workload


def leg_c3() -> dict:
    """120 atoms, 3600 functions, with the kernel choices the autotuner made on the 40-atom
    workload (same shell recipe; tuning here would spend minutes building task tables of the
    larger lane sides): the first FP64 call (Schwarz factors, task tables), one FP64 and one
    mixed build (split C3_SPLIT, after its first call), the energy change of the mixed build, and
    the FP32 tier's own error (bound window)."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    cfg = B.CONFIGS["c3_jk_mixed"]
    mol = B.config_molecule("c3_jk_mixed")
    ctx = lbx.Context(mol)
    applied = 0
    for tier in ("fp64", "fp32"):
        for cls, kernel in TUNED.get(tier, {}).items():
            applied += bool(ctx.set_variant(cls, lbx.VARIANTS.index(kernel)))
    log(f"[bench] c3: {applied} tuned kernel choices of the 40-atom workload applied")
    D = torch.tensor(B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"]),
                     device="cuda")
    J, K = torch.empty_like(D), torch.empty_like(D)
    t0 = time.perf_counter()
    ctx.jk(D, schwarz=SCHWARZ, J=J, K=K)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    log(f"[bench] c3 first FP64 call {first:.1f} s")
    t64, (_, _, st) = timed(lambda: ctx.jk(D, schwarz=SCHWARZ, J=J, K=K), 1, "c3 fp64")
    Jm, Km = torch.empty_like(D), torch.empty_like(D)
    ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED, mixed_thresh=C3_SPLIT, J=Jm, K=Km)
    torch.cuda.synchronize()
    log("[bench] c3 first mixed call done")
    tm, (_, _, sm) = timed(lambda: ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED,
                                          mixed_thresh=C3_SPLIT, J=Jm, K=Km), 1, "c3 mixed")
    dE = energy_difference(D, J, K, Jm, Km)
    E = 0.5 * float((D * J).sum()) - 0.25 * float((D * K).sum())
    res = {"workload": "c3_jk_mixed", "nbf": ctx.nbf, "atoms": mol.natoms,
           "quartets": st["quartets"], "fp64_ms": t64 * 1e3, "mixed_ms": tm * 1e3, "split": C3_SPLIT,
           "speedup_vs_fp64": t64 / tm, "fp32_share": sm["quartets_fp32"] / sm["quartets"],
           "fp64_tflops": st["flops"] / t64 / 1e12, "first_call_s": first,
           "kernel_choices": f"{applied} class-pair kernels as tuned on the 40-atom workload",
           "energy_2e": E, "delta_E": dE,
           "max_abs_dJ": float((Jm - J).abs().max()), "max_abs_dK": float((Km - K).abs().max())}
    del Jm, Km, J, K
    res["fp32_tier"] = fp32_tier_error(ctx, D, per_class=False, split=C3_SPLIT)
    ctx.close()
    log("[bench] c3 done")
    return res


def leg_c4() -> dict:
    """The 256-molecule RI batch: set-up, first call (task tables, graph capture), two replays."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    mols = B.batch_molecules()
    ctx = lbx.Context.batch(mols)
    ctx.set_graphs(True)
    out = torch.empty((len(mols), 4), dtype=torch.float64, device="cuda")
    ctx.ri_sums(out)
    torch.cuda.synchronize()
    t_untuned, _ = timed(lambda: ctx.ri_sums(out), 2, "c4 ri (default kernels)")
    t0 = time.perf_counter()
    ctx.tune()
    t_tune = time.perf_counter() - t0
    ctx.ri_sums(out)
    torch.cuda.synchronize()
    t, (_, st) = timed(lambda: ctx.ri_sums(out), 2, "c4 ri (tuned)")
    ctx.close()
    log("[bench] c4 done")
    return {"workload": "c4_ri_batch", "molecules": len(mols), "ms": t * 1e3,
            "ms_default_kernels": t_untuned * 1e3, "tune_s": t_tune,
            "triplets_per_s": st["triplets"] / t, "doublets_per_s": st["doublets"] / t,
            "tflops": st["flops"] / t / 1e12}


def leg_c5() -> dict:
    """The 60-atom nuclear gradient, one pass: the first call, whose set-up (Schwarz factors,
    E' matrices, task tables) is ~1% of it (67.6 s against 67.5 s for a repeated call), so the
    default run stays within its time budget."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    cfg = B.CONFIGS["c5_gradient"]
    mol = B.config_molecule("c5_gradient")
    ctx = lbx.Context(mol)
    D = torch.tensor(B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"]),
                     device="cuda")
    t, (g, st) = timed(lambda: ctx.jk_grad(D, schwarz=SCHWARZ), 1, "c5 gradient (first call)")
    gh = g.cpu().numpy()
    ctx.close()
    log("[bench] c5 done")
    return {"workload": "c5_gradient", "atoms": mol.natoms, "nbf": int(mol.shells.nbf),
            "ms_first_call": t * 1e3, "quartets": st["quartets"], "quartets_per_s": st["quartets"] / t,
            "tflops": st["flops"] / t / 1e12,
            "max_abs_translation_sum": float(abs(gh.sum(axis=0)).max())}


def leg_ri() -> dict:
    """Density-fitted J/K (lbx_ri_jk) on the 40-atom workload with a 56-function-per-atom
    auxiliary basis, against the exact FP64 build on the same molecule and density."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    cfg = B.CONFIGS["c2_jk_fp64"]
    # the same atoms and main shells as the c2 workload (same seed), plus auxiliary shells
    mol = B.synthetic_molecule(cfg["seed"], cfg["natoms"], cfg["box"], B.RECIPE_MAIN, B.RECIPE_AUX)
    ctx = lbx.Context(mol)
    D = torch.tensor(B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"]),
                     device="cuda")
    ctx.ri_jk(D)
    torch.cuda.synchronize()
    t, (Jr, Kr) = timed(lambda: ctx.ri_jk(D), 2, "ri J/K")
    tj, (Jr2, _) = timed(lambda: ctx.ri_jk(D, exchange=False), 2, "ri J")
    J, K, _ = ctx.jk(D, schwarz=SCHWARZ)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, K, _ = ctx.jk(D, schwarz=SCHWARZ, J=J, K=K)
    torch.cuda.synchronize()
    t_exact = time.perf_counter() - t0
    e_j = 0.5 * float((D * (Jr - J)).sum())
    e_k = -0.25 * float((D * (Kr - K)).sum())
    n, na = ctx.nbf, ctx.nbf_aux
    # dense work of the K part: T_P = W_P D and K = sum_P T_P W_P^T, 2 n^3 naux FLOPs each
    k_flops = 4.0 * n ** 3 * na
    ctx.close()
    log("[bench] ri done")
    return {"workload": "c2 molecule + 56-function auxiliary basis", "nbf": n, "naux": na,
            "ri_jk_ms": t * 1e3, "ri_j_ms": tj * 1e3, "exact_jk_ms": t_exact * 1e3,
            "k_dense_tflops_lower_bound": k_flops / t / 1e12,
            "fit_error_J_energy": e_j, "fit_error_K_energy": e_k}


WATER = ([8, 1, 1], [[0.0, 0.0, 0.0], [0.757, 0.586, 0.0], [-0.757, 0.586, 0.0]])   # angstrom


def leg_scf() -> dict:
    """Restricted Hartree-Fock (lible_amd.scf) of a 2x2x2 grid of water molecules 3.2 angstrom
    apart in def2-SVP, energy and analytic gradient."""
    import numpy as np
    from lible_amd import basis as B
    from lible_amd import scf
    nrs, xyz = water_grid(2)
    path = os.path.join(ROOT, "tests", "golden", "basis", "def2-svp.txt")
    mol = B.molecule_from_basis_files(path, nrs, xyz)
    t0 = time.perf_counter()
    rhf = scf.RHF(mol, nrs)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = rhf.kernel()
    t_scf = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = rhf.gradient()
    t_grad = time.perf_counter() - t0
    rhf.close()
    log("[bench] scf done")
    return {"workload": "(H2O)8 def2-SVP RHF", "atoms": len(nrs), "nbf": int(mol.shells.nbf),
            "energy": r.energy, "converged": r.converged, "iterations": r.iterations,
            "setup_s": t_setup, "scf_s": t_scf, "ms_per_iteration": t_scf / r.iterations * 1e3,
            "gradient_s": t_grad, "max_abs_gradient": float(np.abs(g).max()),
            "max_abs_translation_sum": float(np.abs(g.sum(axis=0)).max())}


def water_grid(n: int):
    """n x n x n water molecules 3.2 angstrom apart: atomic numbers and coordinates (angstrom)."""
    nrs, xyz = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for z, p in zip(*WATER):
                    nrs.append(z)
                    xyz.append([p[0] + 3.2 * i, p[1] + 3.2 * j, p[2] + 3.2 * k])
    return nrs, xyz


def leg_qzvp() -> dict:
    """J/K of (H2O)8 in def2-QZVP (g shells on O: the run-time-L kernels): one build with the
    default kernel choices, the autotuner, one tuned build, the share of the g classes, and one
    J/K nuclear gradient."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    nrs, xyz = water_grid(2)
    path = os.path.join(ROOT, "tests", "golden", "basis", "def2-qzvp.txt")
    mol = B.molecule_from_basis_files(path, nrs, xyz)
    ctx = lbx.Context(mol)
    D = torch.tensor(B.density_from_orbitals(3, ctx.nbf, 40, 0.05), device="cuda")

    def profiled():
        ctx.profile(True)
        ctx.jk(D, schwarz=SCHWARZ)
        prof = ctx.profile_read()
        ctx.profile(False)
        return prof

    ctx.jk(D, schwarz=SCHWARZ)   # first call: Schwarz factors, task tables
    t_default, (_, _, st) = timed(lambda: ctx.jk(D, schwarz=SCHWARZ), 1, "qzvp default")
    p_default = profiled()
    t0 = time.perf_counter()
    ctx.tune(D, schwarz=SCHWARZ)
    t_tune = time.perf_counter() - t0
    t_tuned, _ = timed(lambda: ctx.jk(D, schwarz=SCHWARZ), 1, "qzvp tuned")
    p_tuned = profiled()
    ctx.jk_grad(D, schwarz=SCHWARZ)   # first call: derivative E matrices, task tables
    t_grad, (grad, _) = timed(lambda: ctx.jk_grad(D, schwarz=SCHWARZ), 1, "qzvp gradient")
    nbf = ctx.nbf
    ctx.close()

    def g_share(prof):
        return sum(r["ms"] for r in prof if 4 in r["class"])

    g_flops = sum(r["flops"] for r in p_tuned if 4 in r["class"])
    log("[bench] qzvp done")
    return {"workload": "(H2O)8 def2-QZVP J/K", "nbf": nbf, "quartets": st["quartets"],
            "ms_default_kernels": t_default * 1e3, "ms_g_classes_default": g_share(p_default),
            "tune_s": t_tune, "ms_tuned": t_tuned * 1e3, "ms_g_classes_tuned": g_share(p_tuned),
            "g_classes_tflops": g_flops / max(g_share(p_tuned), 1e-9) / 1e9,
            "ms_gradient": t_grad * 1e3,
            "max_abs_translation_sum": float(grad.sum(dim=0).abs().max()),
            "max_abs_gradient": float(grad.abs().max())}


LEGS = {"c1": leg_c1, "c3": leg_c3, "c4": leg_c4, "c5": leg_c5, "ri": leg_ri, "scf": leg_scf,
        "qzvp": leg_qzvp}


def bench_ri_batch(a):
    """Batched RI (mn|P) + (P|Q) over 256 molecules in one context; a step is one lbx_ri_sums
    (replay of the captured graph): every molecule's eri3 and eri2 evaluated and reduced to its
    checksums."""
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    mols = B.batch_molecules()
    t0 = time.perf_counter()
    ctx = lbx.Context.batch(mols)
    ctx.set_graphs(True)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log(f"[bench] c4_ri_batch: {len(mols)} molecules, {sum(m.natoms for m in mols)} atoms, "
        f"nbf {ctx.nbf}, naux {ctx.nbf_aux}, setup {t_setup:.1f} s")
    out = torch.empty((len(mols), 4), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    _, st = ctx.ri_sums(out)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    log(f"[bench] first call (task tables + graph capture) {t_first:.2f} s")
    for k in range(a.warmup):
        ctx.ri_sums(out)
        log(f"[bench] warm-up step {k} done")
    step_s, (_, st) = timed(lambda: ctx.ri_sums(out), a.steps, "ri")
    ref = out.cpu().numpy().copy()
    # mixed precision: triplets / doublets with Schwarz bound below MIXED_SPLIT in FP32
    outm = torch.empty_like(out)
    ctx.ri_sums(outm, precision=lbx.LBX_MIXED, mixed_thresh=MIXED_SPLIT)
    log("[bench] mixed first call (bounds, FP32 E matrices, graph) done")
    mixed_s, (_, stm) = timed(lambda: ctx.ri_sums(outm, precision=lbx.LBX_MIXED,
                                                  mixed_thresh=MIXED_SPLIT), a.steps, "ri mixed")
    got = outm.cpu().numpy()
    tflops = st["flops"] / step_s / 1e12
    res = {
        "metric": "batched RI eri3 (mn|P) shell triplets/s (+ eri2 doublets/s)",
        "value": st["triplets"] / step_s,
        "unit": "triplets/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "c4_ri_batch", "molecules": len(mols),
                   "atoms_per_molecule": "3..30", "main": "30 AO/atom", "aux": "56 AO/atom",
                   "triplets": st["triplets"], "doublets": st["doublets"]},
        "doublets_per_s": st["doublets"] / step_s,
        "total_tflops": tflops,
        "total_frac_fp64_peak": tflops / FP64_PEAK_TFLOPS,
        "setup_s": t_setup, "first_call_s": t_first,
        "checksum_eri3_abs_total": float(out[:, 0].sum()),
        "mixed": {"split": MIXED_SPLIT, "ms_per_step": mixed_s * 1e3,
                  "triplets_per_s": stm["triplets"] / mixed_s,
                  "fp32_share_triplets": stm["triplets_fp32"] / max(stm["triplets"], 1.0),
                  "fp32_share_doublets": stm["doublets_fp32"] / max(stm["doublets"], 1.0),
                  "speedup_vs_fp64": step_s / mixed_s,
                  "max_rel_err_abs_sums": float((abs(got[:, [0, 2]] - ref[:, [0, 2]]) /
                                                 abs(ref[:, [0, 2]])).max()),
                  "max_abs_err_sums": float(abs(got[:, [1, 3]] - ref[:, [1, 3]]).max())},
    }
    if not a.no_cpu_baseline:
        # the oracle's dense eri3 + eri2 of the first small molecules, ~cpu-seconds of work
        from oracle import oracle as O
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        done_t, done_q, names = 0.0, 0.0, []
        for k, m in enumerate(mols):
            if m.natoms > 8:
                continue
            _, s1 = lbx.Context.batch([m]).ri_sums()
            s = O.Structure(m)
            t0 = time.perf_counter()
            s.eri3()
            s.eri2()
            done_t += time.perf_counter() - t0
            done_q += s1["triplets"]
            names.append(k)
            if done_t > a.cpu_seconds:
                break
        res["cpu_baseline"] = {"value": done_q / max(done_t, 1e-9), "unit": "triplets/s",
                               "cores": cores, "kind": "port",
                               "sample": f"oracle dense eri3 + eri2 of molecules {names} "
                                         f"({done_q:.0f} triplets, {done_t:.1f} s)"}
    print(json.dumps(res), flush=True)
    return 0


def bench_gradient(a):
    """Nuclear gradient of the two-electron energy at fixed density (lbx_jk_grad) on the
    60-atom, 1800-function workload; checked against a central difference of the energy
    along the largest gradient component."""
    import numpy as np
    import torch
    from lible_amd import basis as B
    from lible_amd import lbx
    cfg = B.CONFIGS["c5_gradient"]
    mol = B.config_molecule("c5_gradient")
    ctx = lbx.Context(mol)
    D = torch.tensor(B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"]),
                     device="cuda")
    log(f"[bench] c5_gradient: nbf={ctx.nbf} atoms={mol.natoms}")
    t0 = time.perf_counter()
    g, st = ctx.jk_grad(D, schwarz=SCHWARZ)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    log(f"[bench] first call (Schwarz factors, E' matrices, task tables) {t_first:.2f} s")
    for k in range(a.warmup):
        ctx.jk_grad(D, schwarz=SCHWARZ, grad=g)
        log(f"[bench] warm-up step {k} done")
    step_s, (_, st) = timed(lambda: ctx.jk_grad(D, schwarz=SCHWARZ, grad=g), a.steps, "gradient")
    gh = g.cpu().numpy()
    # central difference along the largest component (h = 1e-4 bohr)
    ia, x = np.unravel_index(np.argmax(np.abs(gh)), gh.shape)
    h = 1e-4
    es = []
    for sgn in (1, -1):
        c = mol.coords.copy()
        c[ia, x] += sgn * h
        cx = lbx.Context(B.Molecule(c, mol.shells))
        J, K, _ = cx.jk(D, schwarz=SCHWARZ)
        es.append(0.5 * float((D * J).sum()) - 0.25 * float((D * K).sum()))
        cx.close()
        log(f"[bench] finite-difference energy {len(es)} done")
    fd = (es[0] - es[1]) / (2 * h)
    # mixed precision: quartets with Schwarz bound below MIXED_SPLIT in FP32
    gm = torch.empty_like(g)
    ctx.jk_grad(D, schwarz=SCHWARZ, grad=gm, precision=lbx.LBX_MIXED, mixed_thresh=MIXED_SPLIT)
    log("[bench] mixed gradient first call done")
    mixed_s, (_, stm) = timed(lambda: ctx.jk_grad(D, schwarz=SCHWARZ, grad=gm, precision=lbx.LBX_MIXED,
                                                  mixed_thresh=MIXED_SPLIT), a.steps, "mixed gradient")
    gmh = gm.cpu().numpy()
    tflops = st["flops"] / step_s / 1e12
    res = {
        "metric": "nuclear gradient of the 2e energy: shell quartets/s",
        "value": st["quartets"] / step_s, "unit": "quartets/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "c5_gradient", "atoms": mol.natoms, "nbf": ctx.nbf,
                   "schwarz": SCHWARZ, "quartets": st["quartets"], "flops_per_step": st["flops"]},
        "total_tflops": tflops, "total_frac_fp64_peak": tflops / FP64_PEAK_TFLOPS,
        "first_call_s": t_first,
        "fd_check": {"atom": int(ia), "axis": int(x), "analytic": float(gh[ia, x]), "central_diff": fd,
                     "abs_diff": abs(float(gh[ia, x]) - fd)},
        "max_abs_translation_sum": float(np.abs(gh.sum(axis=0)).max()),
        "mixed": {"split": MIXED_SPLIT, "ms_per_step": mixed_s * 1e3,
                  "fp32_share": stm["quartets_fp32"] / max(stm["quartets"], 1.0),
                  "speedup_vs_fp64": step_s / mixed_s,
                  "max_abs_dgrad": float(np.abs(gmh - gh).max()),
                  "max_abs_grad": float(np.abs(gh).max())},
    }
    print(json.dumps(res), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2_jk_fp64")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=1,
                    help="timed mixed-precision builds after the FP64 ones (0: skip)")
    ap.add_argument("--legs", default="c1,c3,c4,c5,ri,scf",
                    help="other workloads measured after the main one, one pass each: c1 (dense "
                         "eri4, 84 functions), c3 (120-atom FP64 and mixed J/K), c4 (256-molecule "
                         "RI batch), c5 (60-atom gradient), ri (density-fitted J/K), scf (RHF "
                         "energy and gradient of (H2O)8); also qzvp (J/K of (H2O)8 in def2-QZVP, "
                         "g shells), not in the default list")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="end with a 1 GiB device-to-device copy, the reference reading for "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: the J/K build runs on one GPU; --gpus {a.gpus} is not supported "
              "(J and K would shard by bra-pair blocks with one sum all-reduce, not implemented)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from lible_amd import basis as B
    from lible_amd import lbx

    torch.cuda.set_device(0)
    if a.config == "c4_ri_batch":
        return bench_ri_batch(a)
    if a.config == "c5_gradient":
        return bench_gradient(a)
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
    log(f"[bench] first call (Schwarz factors + task tables) {t_first:.2f} s, "
        f"quartets {st['quartets']:.3e}")
    t0 = time.perf_counter()
    if not a.no_tune:
        ctx.tune(D, schwarz=SCHWARZ)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    log(f"[bench] tuning {t_tune:.2f} s")
    for k in range(a.warmup):
        _, _, st = ctx.jk(D, schwarz=SCHWARZ, J=J, K=K)
        torch.cuda.synchronize()
        log(f"[bench] warm-up step {k} done")

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
    # the tuned kernel per class pair, reused by the 120-atom leg (same shell recipe; tuning
    # there would cost minutes of task-table building for its larger lane sides)
    TUNED["fp64"] = {p["class"]: p["kernel"] for p in prof}

    step_s = elapsed / a.steps
    value = st["quartets"] / step_s
    kern_ms = sum(p["ms"] for p in prof) / a.steps
    dom = max(prof, key=lambda p: p["ms"])
    dom_ms = dom["ms"] / dom["calls"]
    dom_tflops = dom["flops"] / dom["calls"] / (dom_ms * 1e-3) / 1e12
    tot_tflops = st["flops"] / step_s / 1e12
    # energy as a by-product check of the build
    E = 0.5 * float((D * J).sum()) - 0.25 * float((D * K).sum())

    mixed = None
    if a.mixed_steps > 0:
        # the same build with FP32 R and transforms for quartets whose bound is < MIXED_SPLIT
        Jm, Km = torch.empty_like(D), torch.empty_like(D)
        ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED, mixed_thresh=MIXED_SPLIT, J=Jm, K=Km)
        torch.cuda.synchronize()
        log("[bench] first mixed call (FP32 task tables) done")
        ctx.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(a.mixed_steps):
            _, _, sm = ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED,
                              mixed_thresh=MIXED_SPLIT, J=Jm, K=Km)
            log(f"[bench] mixed step {k} done")
        torch.cuda.synchronize()
        m_s = (time.perf_counter() - t0) / a.mixed_steps
        TUNED["fp32"] = {p["class"]: p["kernel"] for p in ctx.profile_read() if p["kernel"].endswith("f32")}
        ctx.profile(False)
        dE = energy_difference(D, J, K, Jm, Km)
        tier = fp32_tier_error(ctx, D, per_class=True)
        log("[bench] FP32-tier error per class pair done")
        mixed = {
            "split": MIXED_SPLIT,
            "ms_per_step": m_s * 1e3,
            "quartets_per_s": sm["quartets"] / m_s,
            "fp32_share": sm["quartets_fp32"] / max(sm["quartets"], 1),
            "speedup_vs_fp64": step_s / m_s,
            "delta_E": dE,
            "delta_E_rel": dE / abs(E),
            "max_abs_dJ": float((Jm - J).abs().max()),
            "max_abs_dK": float((Km - K).abs().max()),
            "fp32_tier": tier,
        }
        del Jm, Km

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
            # FP64 compute bound (MI355X's 78.6 TFLOP/s FP64 peak is the same for the matrix and
            # the vector units); compute_unit says which of them the kernel's transforms use
            "bound": "mfma",
            "compute_unit": "valu" if dom["kernel"].startswith("valu") else "mfma",
            "kernel": kernel_name(dom),
            "achieved": dom_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": dom_tflops / FP64_PEAK_TFLOPS,
            # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
            # this kernel (not measured in this run; null when they cover another kernel)
            "traffic": measured_traffic(kernel_name(dom)),
            "traffic_source": "profiles/jk_c2_traffic.json (separate rocprofv3 counter runs)",
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
        "mixed": mixed,
    }
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline (oracle sample)")
        out["cpu_baseline"] = cpu_baseline(mol, Dh, st, a.cpu_seconds)
    # the other workloads, one measured pass each, so the driver's run records them too
    ctx.close()
    del ctx, D, J, K
    torch.cuda.empty_cache()
    legs = {}
    for leg in [x for x in a.legs.split(",") if x]:
        legs[leg] = LEGS[leg]()
        torch.cuda.empty_cache()
    out["legs"] = legs
    print(json.dumps(out), flush=True)
    if a.pmc_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy done: 1 GiB read, 1 GiB written")
    return 0


if __name__ == "__main__":
    sys.exit(main())
