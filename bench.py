"""Throughput of the octree FV curl-curl forward solve on one MI355X.

Workload (one step): the seeded random 3-D model (3 log-uniform layers plus
1-4 blocks, octree levels 3-7, 11x11 stations at 1 km, 8 frequencies
log-spaced 1e-3..1e2 Hz x 2 polarisations = 16 systems, complex128), solved to
relative residual 1e-7 with the reference's forward path (scattered field,
SSOR-BiCGStab from zero, dead-edge interpolation, post-hoc DivCorr) as one
batched device call whose inputs (rhs_total, E_pri) already sit in HBM.

value = solved systems per second over the K timed steps.  The JSON line also
carries the roofline of the batched SpMV kernel (HIP events on the library's
stream), the mixed-precision variant, and the CPU oracle timed on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--half-width", type=float, default=100e3)
    ap.add_argument("--station-radius", type=float, default=80e3)
    ap.add_argument("--max-level", type=int, default=7)
    ap.add_argument("--maxiter", type=int, default=40000,
                    help="BiCGStab iteration cap per system (the 1e-3 Hz systems of the plain "
                         "FP64 method need 12-25 thousand)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work")
    ap.add_argument("--kernel-only", action="store_true",
                    help="profiling mode: build the workload, then only run the batched SpMV "
                         "and SSOR kernels --reps times and a 1 GiB device copy (counter "
                         "calibration)")
    ap.add_argument("--reps", type=int, default=20)
    return ap.parse_args()


def kernel_only(args):
    """Launch only the roofline kernels (for rocprofv3 --pmc runs)."""
    import torch

    from maple3dmt_amd import Context, configs
    wl = configs.random3d(args.seed, max_level=args.max_level, half_width=args.half_width,
                          station_radius=args.station_radius)
    ctx = Context(0)
    model = configs.build_model(wl, ctx)
    omegas = np.repeat(2 * np.pi * wl.freqs, 2)
    ms, by = model.device.bench_kernel(0, omegas, reps=args.reps)
    ms2, by2 = model.device.bench_kernel(1, omegas, reps=max(1, args.reps // 4))
    iters = {}
    for kind, tag in ((2, "fp64"), (3, "c64")):
        for k in (1, 2, 4, 16):
            t, b = model.device.bench_kernel(kind, omegas[:k], reps=max(1, args.reps // 4))
            iters[f"{tag}_ns{k}"] = {"ms": t, "alg_GBs": b / t / 1e6}
            log(f"[kernel-only] BiCGStab iteration {tag} x{k}: {t:.3f} ms, {b / t / 1e6:.0f} GB/s")
    x = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"spmv_ms": ms, "spmv_alg_bytes": by, "ssor_ms": ms2, "ssor_alg_bytes": by2,
                      "iteration": iters,
                      "copy_bytes_each_way": x.numel() * 8, "copies": 4, "edges": model.n}),
          flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.kernel_only:
        kernel_only(args)
        return

    import torch

    from maple3dmt_amd import MODE_FORWARD, PREC_MIXED, Context, configs

    t0 = time.time()
    wl = configs.random3d(args.seed, max_level=args.max_level, half_width=args.half_width,
                          station_radius=args.station_radius, maxiter=args.maxiter)
    ctx = Context(0)
    model = configs.build_model(wl, ctx)
    op = model.device
    n = model.n
    log(f"[bench] mesh: {model.nc} cells, {n} edges, {model.nn} nodes, "
        f"{model.count('hanging_edges')} dead edges, built in {time.time() - t0:.1f}s")

    nf = len(wl.freqs)
    ns = 2 * nf
    omegas = np.repeat(2 * np.pi * wl.freqs, 2)
    rhs = np.zeros((ns, n), np.complex128)
    aux = np.zeros((ns, n), np.complex128)
    for f, fr in enumerate(wl.freqs):
        w = 2 * np.pi * fr
        for pol in range(2):
            rhs[2 * f + pol] = model.rhs_total(model.bc_values(w, pol))
            aux[2 * f + pol] = model.primary_field(w, pol)
    dev = torch.device("cuda:0")
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_aux = torch.from_numpy(aux).to(dev)
    d_x = torch.zeros_like(d_rhs)
    torch.cuda.synchronize()

    prm64 = wl.fwd.solve_params()

    def step(prm):
        return op.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(),
                               d_x.data_ptr(), ns, prm)

    for _ in range(args.warmup):
        res = step(prm64)
        log(f"[bench] warmup: iterations {[r['iterations'] for r in res]}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        res = step(prm64)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    ms_per_step = 1e3 * dt / args.steps
    value = ns * args.steps / dt
    iters = [r["iterations"] for r in res]
    conv = all(r["converged"] for r in res)
    log(f"[bench] fp64: {ms_per_step:.1f} ms/step, {value:.2f} solves/s, converged={conv}, "
        f"true residual max {max(r['true_residual'] for r in res):.2e}")
    x64 = d_x.cpu().numpy()
    z64 = np.stack([model.extract_impedance(omegas[2 * f], x64[2 * f], x64[2 * f + 1])
                    for f in range(nf)])

    # Mixed precision: complex64 inner BiCGStab inside FP64 refinement.
    mixed = None
    if not args.no_mixed:
        prm32 = wl.fwd.solve_params(precision=PREC_MIXED, inner_tol=1e-3)
        step(prm32)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            r32 = step(prm32)
        torch.cuda.synchronize()
        d32 = time.perf_counter() - t2
        x32 = d_x.cpu().numpy()
        z32 = np.stack([model.extract_impedance(omegas[2 * f], x32[2 * f], x32[2 * f + 1])
                        for f in range(nf)])
        off = (slice(None), slice(None), slice(1, 3))
        dz = float(np.max(np.abs(z32[off] - z64[off]) / np.abs(z64[off])))
        mixed = {"solves_per_s": ns * args.steps / d32, "ms_per_step": 1e3 * d32 / args.steps,
                 "converged": all(r["converged"] for r in r32),
                 "max_rel_dZ_vs_fp64": dz,
                 "iterations": [r["iterations"] for r in r32]}
        log(f"[bench] mixed: {mixed}")

    # Roofline of the batched SpMV (the operator apply of every iteration).
    ms_spmv, bytes_spmv = op.bench_kernel(0, omegas, reps=20)
    ms_ssor, bytes_ssor = op.bench_kernel(1, omegas, reps=5)
    ach = bytes_spmv / (ms_spmv * 1e-3) / 1e9
    info = op.info()
    log(f"[bench] spmv {ms_spmv:.3f} ms, {ach:.0f} GB/s; ssor {ms_ssor:.3f} ms, "
        f"{bytes_ssor / (ms_ssor * 1e-3) / 1e9:.0f} GB/s; {info}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(model, wl, args.cpu_budget, iters)

    line = {
        "metric": "MT solves/s at rel-resid 1e-7 (plain FP64 vs batched mixed+learned); SpMV TB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (f64)",
        "data": "synthetic (seeded random 3-D model)",
        "config": {"workload": wl.name + "_11x11st_8f_2pol_fwd", "systems_per_step": ns,
                   "edges": n, "cells": model.nc, "nodes": model.nn, "tol": wl.fwd.bicgstab_tol,
                   "precision": "fp64", "batched": True},
        "converged": conv,
        "iterations": iters,
        "roofline": {"bound": "hbm", "kernel": "batched SpMV (k_spmv, 16 systems)",
                     "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": bytes_spmv, "avg_launch_ms": ms_spmv},
        "ssor_apply": {"avg_launch_ms": ms_ssor, "alg_bytes": bytes_ssor,
                       "achieved_GBs": bytes_ssor / (ms_ssor * 1e-3) / 1e9},
        "spmv_TBs": ach / 1e3,
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    model.close()
    ctx.close()


def cpu_baseline(model, wl, budget, gpu_iters):
    """The CPU oracle (natural-order SSOR-BiCGStab, complex128, one core) on a
    bounded sample: both polarisations of the highest frequency solved to
    completion if that fits the budget, otherwise the measured time per
    iteration scaled by the GPU's iteration counts."""
    from oracle.oracle import Oracle
    t = time.perf_counter()
    o = Oracle(model, sigma_bg=wl.sigma_bg)
    o.set_sigma(model.array("sigma"))
    setup = time.perf_counter() - t
    f = float(wl.freqs[-1])
    o.assemble(2 * np.pi * f)
    # Time per BiCGStab iteration from a capped solve.
    ptr, col, val = o.csr("A")
    rhs, epri = o.rhs(0)
    b = rhs - o.spmv(epri)
    from oracle.oracle import bicgstab_csr
    t = time.perf_counter()
    _, r = bicgstab_csr(ptr, col, val, b, tol=1e-7, maxiter=20)
    per_iter = (time.perf_counter() - t) / max(r["iterations"], 1)
    est_solve = per_iter * float(np.mean(gpu_iters))
    sample = (f"oracle setup {setup:.1f}s; 20 SSOR-BiCGStab iterations on the {f:g} Hz system "
              f"({per_iter * 1e3:.0f} ms/iteration), scaled by the mean GPU iteration count "
              f"{np.mean(gpu_iters):.0f} per system")
    return {"value": 1.0 / est_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": sample, "s_per_iteration": per_iter}


if __name__ == "__main__":
    main()
