"""Throughput of the octree FV curl-curl forward solve on one MI355X.

Workload (one step): the seeded random 3-D model (3 log-uniform layers plus
1-4 blocks, octree levels 3-7, 11x11 stations at 1 km, 8 frequencies
log-spaced 1e-3..1e2 Hz x 2 polarisations = 16 systems, complex128), solved to
relative residual 1e-7 with the reference's forward path (scattered field,
SSOR-BiCGStab from zero, dead-edge interpolation, post-hoc DivCorr) as one
batched device call whose inputs (rhs_total, E_pri) already sit in HBM.

value = solved systems per second of the plain FP64 method (the reference's
forward solve, batched) over the K timed steps.  The JSON line also carries,
under "variants", the same batch solved
  mixed                      with complex64 inner solves in FP64 refinement
  fp64_interleaved           with E_sec by BiCGStab bursts of 100 iterations
                             each followed by DivCorr (solve_with_divcorr_)
  mixed_interleaved          both: the fast path
  mixed_interleaved_tol1e-9  the fast path at relative residual 1e-9
  mixed_interleaved_dc32     the fast path with its DivCorr PCG between bursts
                             in FP32 (M3D_DIVCORR_FP32=1; FP64 by default)
  mixed_interleaved_learned  the fast path started from the trained network's
                             e0 (inference and prolongation inside the timed
                             region), when its weights exist
each with its speed-up over the plain FP64 batch and its largest |dZ|/|Z|
against it; and
  gradient       one misfit gradient of the bench model (fast path): full fields
                 through the host against the device-resident path
  sigma_update   one line-search sigma update: host recompute + upload against
                 the recompute on the device (m3d_operator_update_sigma)
  launch_shapes  the autotuner's launch shape per batch width and precision
                 (tuned on the bench mesh before any timed region)
  roofline       the kernel that dominates the timed region (the backward SSOR
                 sweep at the tail batch width), timed with HIP events on the
                 library's stream, against HBM peak
  spmv           the batched SpMV of all 16 systems (the metric's SpMV TB/s)
  multi_model    misfit gradients (forward + adjoint systems) of several models
                 on meshes of uneven size, sequential and concurrent
  network        the e0 network's BF16 MFMA kernels: inference and one training
                 step on a batch of 8 samples, achieved TFLOP/s against the
                 dense BF16 peak, plus the rocprofv3 MfmaUtil / MfmaFlopsBF16
                 counters of the committed profile
  cpu_baseline   the CPU oracle timed on a bounded sample of the same workload,
                 on one core and on P cores at once (value: the P-core rate)

python bench.py --c3 instead prints one line for the 16-model batch (1024
forward and adjoint systems on meshes of 0.24-2.4 M edges, four streams).
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
MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16
WEIGHTS = os.path.join(ROOT, "maple3dmt_amd", "data", "e0_net.npz")
# HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes,
# calibrated on a 1 GiB device copy (see the summary beside it).
TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
# Per-kernel MFMA counters of the network (rocprofv3 --pmc MfmaUtil MfmaFlopsBF16).
MFMA = os.path.join(ROOT, "profiles", "mfma_e0_network.json")


def measured_traffic(kernel: str, field: str = "traffic_bytes"):
    """HBM bytes per launch of `kernel` (or another field of its entry, e.g.
    the rocprofv3 mean launch time) from the committed counter summary; None
    when it holds no entry for the kernel."""
    try:
        with open(TRAFFIC) as f:
            k = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None or field not in k else float(k[field])


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
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
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--models", type=int, default=4,
                    help="models of the multi-model gradient leg (0 skips it)")
    ap.add_argument("--dc-interval", type=int, default=100,
                    help="BiCGStab iterations between DivCorr steps in the interleaved variants")
    ap.add_argument("--tail-width", type=int, default=2,
                    help="batch width of the roofline kernel (the width the low-frequency tail runs at)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work")
    ap.add_argument("--kernel-only", action="store_true",
                    help="profiling mode: build the workload, then only run the benchmark "
                         "kernels (SpMV, SSOR, BiCGStab iterations, backward sweep) and four "
                         "1 GiB device copies (counter calibration)")
    ap.add_argument("--c3", action="store_true",
                    help="separate mode: the 16-model batch (1024 forward + adjoint systems on "
                         "meshes of 0.24-2.4 M edges), one JSON line of its own")
    ap.add_argument("--c3-precision", choices=("fp64", "mixed"), default="fp64",
                    help="inner precision of the 16-model batch (--c3)")
    ap.add_argument("--fast-only", action="store_true",
                    help="profiling mode: build the workload, tune the launch shapes, then only solve "
                         "the batch twice with the fast path (mixed precision + interleaved DivCorr)")
    ap.add_argument("--network-only", action="store_true",
                    help="profiling mode: only the e0 network's inference and training steps")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of tuning them on the bench "
                         "mesh before the timed region")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_only(args):
    """Launch only the benchmark kernels (for rocprofv3 --pmc runs)."""
    import torch

    from maple3dmt_amd import Context, configs
    wl = configs.random3d(args.seed, max_level=args.max_level, half_width=args.half_width,
                          station_radius=args.station_radius)
    ctx = Context(0)
    model = configs.build_model(wl, ctx)
    omegas = np.repeat(2 * np.pi * wl.freqs, 2)
    dev = model.device
    # The 16-system SpMV in both block layouts, separate blocks first, then
    # packed (the counter passes tell the two apart by launch order).
    base = dev.get_shape(0, 16)
    # Then the LDS-staged row tiles (the default at 16 systems).
    spmv = {}
    for name, packed, tiled in (("packed0", 0, 0), ("packed1", 1, 0), ("tiled", 1, 1)):
        dev.set_shape(0, 16, base["sweep_sub"], base["sweep_cap"], base["row_cap"], base["sweep_chunk"], packed,
                      tiled=tiled)
        t, b = dev.bench_kernel(0, omegas, reps=args.reps)
        spmv[name] = {"ms": t, "alg_bytes": b}
        log(f"[kernel-only] SpMV x16 {name}: {t:.3f} ms, {b / t / 1e6:.0f} GB/s")
    dev.set_shape(0, 16, **base)
    ms, by = spmv["tiled"]["ms"], spmv["tiled"]["alg_bytes"]
    ms2, by2 = model.device.bench_kernel(1, omegas, reps=max(1, args.reps // 4))
    ms4, by4 = model.device.bench_kernel(4, omegas[:args.tail_width], reps=args.reps)
    iters = {}
    for kind, tag in ((2, "fp64"), (3, "c64")):
        for k in (1, 2, 4, 8, 16):
            t, b = model.device.bench_kernel(kind, omegas[:k], reps=max(1, args.reps // 4))
            iters[f"{tag}_ns{k}"] = {"ms": t, "alg_GBs": b / t / 1e6}
            log(f"[kernel-only] BiCGStab iteration {tag} x{k}: {t:.3f} ms, {b / t / 1e6:.0f} GB/s")
    x = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"spmv_ms": ms, "spmv_alg_bytes": by, "spmv_layouts": spmv, "ssor_ms": ms2, "ssor_alg_bytes": by2,
                      "sweep_bwd_ms": ms4, "sweep_bwd_alg_bytes": by4, "iteration": iters,
                      "copy_bytes_each_way": x.numel() * 8, "copies": 4, "edges": model.n}),
          flush=True)


def fast_only(args):
    """Only the fast-path solves of the bench batch (for rocprofv3 runs)."""
    import torch

    from maple3dmt_amd import MODE_FORWARD, Context, configs
    wl = configs.random3d(args.seed, max_level=args.max_level, half_width=args.half_width,
                          station_radius=args.station_radius, maxiter=args.maxiter)
    ctx = Context(0)
    model = configs.build_model(wl, ctx)
    om = np.repeat(2 * np.pi * wl.freqs, 2)
    if not args.no_autotune:
        model.device.autotune(np.resize(om, 32), widths=(1, 2, 4, 8, 16), reps=16)
    rhs = np.stack([model.rhs_total(model.bc_values(w, p % 2)) for p, w in enumerate(om)])
    aux = np.stack([model.primary_field(w, p % 2) for p, w in enumerate(om)])
    dev = torch.device("cuda:0")
    d_rhs, d_aux = torch.from_numpy(rhs).to(dev), torch.from_numpy(aux).to(dev)
    prm = wl.fwd.fast_params(dc_interval=args.dc_interval)
    for _ in range(2):
        d_x = torch.zeros_like(d_rhs)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = model.device.solve_device(MODE_FORWARD, om, d_rhs.data_ptr(), d_aux.data_ptr(), d_x.data_ptr(),
                                      len(om), prm)
        log(f"[fast-only] {time.perf_counter() - t:.3f}s, iterations {[q['iterations'] for q in r]}")
    model.close()
    ctx.close()


def impedances(model, omegas, x, nf):
    return np.stack([model.extract_impedance(omegas[2 * f], x[2 * f], x[2 * f + 1]) for f in range(nf)])


def rel_dz(z, zref):
    off = (slice(None), slice(None), slice(1, 3))
    return float(np.max(np.abs(z[off] - zref[off]) / np.abs(zref[off])))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.kernel_only:
        kernel_only(args)
        return
    if args.fast_only:
        fast_only(args)
        return
    if args.c3:
        print(json.dumps(c3_leg(args)), flush=True)
        return
    if args.network_only:
        from maple3dmt_amd import Context
        ctx = Context(0)
        print(json.dumps(network_leg(args, ctx)), flush=True)
        ctx.close()
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
    shapes = None
    if not args.no_autotune:
        # Launch shapes per batch width and precision, timed on this mesh
        # (outside every timed region).
        t1 = time.time()
        tuned = op.autotune(np.resize(omegas, 32), widths=(1, 2, 4, 8, 16), reps=16)
        shapes = {f"{'c64' if p else 'fp64'}_x{w}": {"ms_per_iteration": round(ms, 4), **sh}
                  for (p, w), (ms, sh) in tuned.items()}
        log(f"[bench] autotune {time.time() - t1:.1f}s: {shapes}")
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

    def solve(prm, x):
        return op.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(), x.data_ptr(),
                               ns, prm)

    def timed(prm, make_x, label):
        """Warm-up, then K timed steps; make_x() runs inside the timed region."""
        for _ in range(args.warmup):
            res = solve(prm, make_x())
            log(f"[bench] {label} warmup: iterations {[r['iterations'] for r in res]}")
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            x = make_x()
            res = solve(prm, x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return res, x, dt

    # Plain FP64 (the reference's method, batched).
    prm64 = wl.fwd.solve_params()

    def zero_x():
        d_x.zero_()
        torch.cuda.synchronize()  # the library runs on its own stream
        return d_x

    res, x64, dt = timed(prm64, zero_x, "fp64")
    ms_per_step = 1e3 * dt / args.steps
    value = ns * args.steps / dt
    iters = [r["iterations"] for r in res]
    conv = all(r["converged"] for r in res)
    log(f"[bench] fp64: {ms_per_step:.1f} ms/step, {value:.2f} solves/s, converged={conv}, "
        f"iterations {iters}, true residual max {max(r['true_residual'] for r in res):.2e}")
    z64 = impedances(model, omegas, x64.cpu().numpy(), nf)
    # Converged reference Z (untimed): the fast path to relative residual
    # 1e-10.  Every leg reports its largest |dZ|/|Z| against it, the error
    # its tolerance leaves.
    t_ref = time.perf_counter()
    ref_res = solve(wl.fwd.fast_params(dc_interval=args.dc_interval, tol=1e-10), zero_x())
    z_conv = impedances(model, omegas, d_x.cpu().numpy(), nf)
    log(f"[bench] converged reference (tol 1e-10) in {time.perf_counter() - t_ref:.1f}s, "
        f"converged {all(r['converged'] for r in ref_res)}")
    fp64_dz_conv = rel_dz(z64, z_conv)

    def leg(label, prm, make_x=zero_x):
        r, x, d = timed(prm, make_x, label)
        zx = impedances(model, omegas, x.cpu().numpy(), nf)
        out = {"solves_per_s": ns * args.steps / d, "ms_per_step": 1e3 * d / args.steps,
               "speedup_vs_fp64": (ns * args.steps / d) / value,
               "tol": prm.tol,
               "converged": all(q["converged"] for q in r),
               "max_rel_dZ_vs_fp64": rel_dz(zx, z64),
               "max_rel_dZ_vs_converged": rel_dz(zx, z_conv),
               "iterations": [q["iterations"] for q in r]}
        if prm.use_initial_guess:
            out["guess_used"] = [int(q["guess_used"]) for q in r]
            out["guess_residual"] = [round(q["guess_residual"], 4) for q in r]
        log(f"[bench] {label}: {out}")
        return out

    def params(precision, interleaved, tol=None):
        if interleaved:
            # E_sec by BiCGStab bursts of 100 iterations (the interval of the
            # reference's configuration 1), each followed by DivCorr.
            return wl.fwd.fast_params(precision=precision, dc_interval=args.dc_interval, tol=tol)
        return wl.fwd.solve_params(precision=precision, inner_tol=1e-3 if precision else 0.0, tol=tol)

    legs = {}
    if not args.no_mixed:
        # Precision alone: complex64 inner BiCGStab inside FP64 refinement.
        legs["mixed"] = leg("mixed", params(PREC_MIXED, False))
        # Interleaved DivCorr alone, FP64.
        legs["fp64_interleaved"] = leg("fp64+interleaved", params(0, True))
        # The fast path: mixed precision + interleaved DivCorr.
        legs["mixed_interleaved"] = leg("mixed+interleaved", params(PREC_MIXED, True))
        # The fast path at relative residual 1e-9, the tolerance at which every
        # path's Z is converged to 1e-6 on the level-6 parity models (the
        # reference's 1e-7 leaves up to 1e-4 there at 19-100 Hz).
        legs["mixed_interleaved_tol1e-9"] = leg("mixed+interleaved tol 1e-9", params(PREC_MIXED, True, 1e-9))
        # The fast path with the DivCorr PCG between bursts in FP32 (opt-in:
        # the specified precision split keeps the potential solve in FP64).
        saved = os.environ.get("M3D_DIVCORR_FP32")
        os.environ["M3D_DIVCORR_FP32"] = "1"
        try:
            legs["mixed_interleaved_dc32"] = leg("mixed+interleaved, FP32 DivCorr PCG", params(PREC_MIXED, True))
        finally:
            if saved is None:
                del os.environ["M3D_DIVCORR_FP32"]
            else:
                os.environ["M3D_DIVCORR_FP32"] = saved
        if not args.no_learned and os.path.exists(args.weights):
            from maple3dmt_amd.learned import Box, CoarseNet, initial_guess
            wts = np.load(args.weights)
            box = Box(*wts["box"].tolist())
            net = CoarseNet(ctx, max_batch=nf, dilations=tuple(int(v) for v in wts["dilations"]))
            net.set_params(wts["w"], wts["b"])
            prml = params(PREC_MIXED, True)
            prml.use_initial_guess = 2
            legs["mixed_interleaved_learned"] = leg("mixed+interleaved+learned", prml,
                                                    lambda: initial_guess(net, model, wl, box))
            legs["mixed_interleaved_learned"]["weights"] = os.path.relpath(args.weights, ROOT)
            net.close()

    # Roofline of the kernel that dominates the timed region: the backward
    # SSOR sweep at the tail width (the low-frequency systems run their
    # thousands of iterations there); and the 16-system batched SpMV.
    tw = max(1, min(args.tail_width, ns))
    sub = op.get_shape(0, 1 << (tw - 1).bit_length())["sweep_sub"]
    sweep_name = f"k_sweep_bwd<{tw}, double, {sub}>"
    # Algorithmic bytes count a coupling as its FP64 value and int32 column;
    # the *_stored_format figures use the 4-byte word the packed blocks and
    # row tiles actually hold.
    ms_sw, bytes_sw, stored_sw = op.bench_kernel(4, omegas[:tw], reps=args.reps, stored=True)
    ach = bytes_sw / (ms_sw * 1e-3) / 1e9
    ms_spmv, bytes_spmv, stored_spmv = op.bench_kernel(0, omegas, reps=args.reps, stored=True)
    ach_spmv = bytes_spmv / (ms_spmv * 1e-3) / 1e9
    sh16 = op.get_shape(0, 16)
    spmv_kernel = (f"k_spmv_tiled<{ns}, double>" if sh16["tiled"]
                   else f"k_spmv<{ns}, double> packed{sh16['packed']}")
    # The same product gathering x from device memory (SELL blocks), for comparison.
    op.set_shape(0, 16, **{**sh16, "tiled": 0})
    ms_sell, _ = op.bench_kernel(0, omegas, reps=args.reps)
    op.set_shape(0, 16, **sh16)
    ms_ssor, bytes_ssor = op.bench_kernel(1, omegas, reps=max(1, args.reps // 4))
    info = op.info()
    info["device_bytes"] = op.device_bytes()
    log(f"[bench] sweep_bwd x{tw} {ms_sw * 1e3:.1f} us/launch, {ach:.0f} GB/s; spmv x{ns} {ms_spmv:.3f} ms, "
        f"{ach_spmv:.0f} GB/s; ssor x{ns} {ms_ssor:.3f} ms; {info}")

    # A line-search trial's sigma update: host recompute of Me and the DivCorr
    # Laplacian plus upload, against the recompute on the device.
    s0 = model.array("sigma").copy()
    s1 = s0 * np.exp(np.random.default_rng(5).uniform(-0.5, 0.5, s0.shape))
    t = time.perf_counter()
    model.set_sigma(s1)
    host_ms = 1e3 * (time.perf_counter() - t)
    model.set_sigma(s0)
    t = time.perf_counter()
    model.update_sigma_device(s1)
    dev_ms = 1e3 * (time.perf_counter() - t)
    model.set_sigma(s0)
    sigma_update = {"cells": model.nc, "host_ms": host_ms, "device_ms": dev_ms}
    log(f"[bench] sigma update: host {host_ms:.1f} ms, device {dev_ms:.2f} ms")

    # Misfit gradient of the bench model (16 forward + 16 adjoint systems,
    # fast path): full fields through the host against the device-resident
    # path that moves only the station cells' edges.
    gradient = None
    if not args.no_mixed:
        from maple3dmt_amd import inversion
        from maple3dmt_amd.multimodel import halfspace_data
        obs, err = halfspace_data(wl)
        gp = wl.fwd.fast_params(dc_interval=args.dc_interval)
        inversion.gradient_device(model, wl.freqs, obs, err, gp, gp)      # warm-up
        t = time.perf_counter()
        g_host = inversion.gradient(model, wl.freqs, obs, err, gp, gp)
        t_host = time.perf_counter() - t
        t = time.perf_counter()
        g_dev = inversion.gradient_device(model, wl.freqs, obs, err, gp, gp)
        t_dev = time.perf_counter() - t
        gradient = {"systems": 4 * nf, "host_fields_s": t_host, "device_resident_s": t_dev,
                    "station_edges": int(len(model.station_edges())),
                    "max_rel_diff": float(np.abs(g_dev.grad - g_host.grad).max() / np.abs(g_host.grad).max())}
        log(f"[bench] gradient: {gradient}")

    models = None
    if args.models > 0:
        models = multi_model_leg(args, model, ctx)

    network = network_leg(args, ctx)

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
                   "maxiter": args.maxiter, "precision": "fp64", "batched": True,
                   # The reference's default cap (forward_solver_fv.h:37) would stop
                   # these systems unconverged; the bench raises it.
                   "reference_default_maxiter": 2000,
                   "systems_over_reference_maxiter": int(sum(i > 2000 for i in iters))},
        "converged": conv,
        "iterations": iters,
        "max_rel_dZ_vs_converged": fp64_dz_conv,
        "roofline": {"bound": "hbm",
                     "kernel": f"{sweep_name} (backward SSOR sweep, one colour per launch)",
                     "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic(sweep_name),
                     "alg_bytes_per_launch": bytes_sw, "avg_launch_ms": ms_sw,
                     "alg_bytes_stored_format_per_launch": stored_sw,
                     "frac_stored_format": stored_sw / (ms_sw * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     # The same launches in the committed kernel trace: rocprofv3
                     # waits on every dispatch's completion signal, about 1 us more
                     # per short launch than back to back in a graph.
                     "rocprof_avg_launch_ms": (lambda v: None if v is None else v / 1e3)(
                         measured_traffic(sweep_name, "rocprof_avg_launch_us"))},
        "spmv": {"kernel": spmv_kernel, "avg_launch_ms": ms_spmv, "alg_bytes": bytes_spmv,
                 "achieved_GBs": ach_spmv, "TBs": ach_spmv / 1e3, "frac": ach_spmv / HBM_PEAK_GBS,
                 "alg_bytes_stored_format": stored_spmv,
                 "TBs_stored_format": stored_spmv / (ms_spmv * 1e-3) / 1e12,
                 "traffic": measured_traffic(spmv_kernel),
                 "rocprof_avg_launch_ms": (lambda v: None if v is None else v / 1e3)(
                     measured_traffic(spmv_kernel, "rocprof_avg_launch_us")),
                 "sell_gather_avg_launch_ms": ms_sell},
        "ssor_apply": {"systems": ns, "avg_ms": ms_ssor, "alg_bytes": bytes_ssor,
                       "achieved_GBs": bytes_ssor / (ms_ssor * 1e-3) / 1e9},
        "operator": info,
        "launch_shapes": shapes,
        "sigma_update": sigma_update,
        "gradient": gradient,
        "variants": legs,
        "multi_model": models,
        "network": network,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    model.close()
    ctx.close()


def multi_model_leg(args, bench_model, ctx):
    """Misfit gradients of several seeded random models on meshes of uneven
    size (octree levels 5-7), each with 16 frequencies x 2 polarisations of
    forward and adjoint systems, by the fast path (mixed precision +
    interleaved DivCorr): one after another, then concurrently on one stream
    per model (multimodel.py).  Reports systems/s of both."""
    from maple3dmt_amd import PREC_MIXED, Context, configs, multimodel
    levels = [5, 6, 7, 6, 5, 7, 6, 5][:args.models]

    def jobs():
        out = []
        for i, lev in enumerate(levels):
            wl = configs.random3d(100 + i, max_level=lev, n_freq=16, half_width=100e3,
                                  station_radius=20e3, maxiter=args.maxiter)
            z, err = multimodel.halfspace_data(wl)
            out.append(multimodel.ModelJob(wl, z, err))
        return out

    prm = configs.ForwardParams().fast_params(dc_interval=args.dc_interval)
    prm.maxiter = args.maxiter
    contexts = [Context(0) for _ in range(len(levels))]
    t = time.time()
    js = jobs()
    multimodel.prepare(js, contexts)
    log(f"[bench] multi-model: {len(levels)} models built in {time.time() - t:.1f}s, "
        f"edges {[j.model.n for j in js]}")
    multimodel.run(js, 1, prm, prm)                # warm-up: workspaces and graphs
    t_seq = multimodel.run(js, 1, prm, prm)
    phi_seq = [j.result.phi for j in js]
    t_par = multimodel.run(js, len(levels), prm, prm)
    n_sys = multimodel.systems(js)
    same = all(a == j.result.phi for a, j in zip(phi_seq, js))
    seq = js
    out = {"models": len(levels), "levels": levels, "edges": [j.model.n for j in seq],
           "systems": n_sys, "frequencies": 16,
           "method": "mixed precision + interleaved DivCorr, forward then adjoint per model",
           "sequential_s": t_seq, "concurrent_s": t_par,
           "systems_per_s_sequential": n_sys / t_seq, "systems_per_s_concurrent": n_sys / t_par,
           "concurrent_matches_sequential": same}
    log(f"[bench] multi-model: {out}")
    for j in js:
        j.model.close()
    for c in contexts:
        c.close()
    return out


def c3_leg(args):
    """Misfit gradients of 16 seeded random models on their own meshes of
    uneven size (octree levels 6-7, 0.24-2.4 M edges), 16 frequencies
    log-spaced 1e-3..1e2 Hz x 2 polarisations, forward and adjoint: 1024
    complex128 systems, solved by BiCGStab with interleaved DivCorr (the
    reference's adjoint method, bursts of 100 iterations, DivCorr PCG to
    1e-4 inside the bursts) in batches of 32 per model and phase, four
    models at a time on four streams.  The models are built on a context of
    their own and taken, largest first, from one queue by four workers, each
    moving the model it takes to its stream.  Each model is released once
    its gradient is done, so at most four models' solve workspaces are
    resident."""
    from maple3dmt_amd import PREC_MIXED, Context, configs, multimodel
    sizes = [(6, 80e3), (7, 40e3), (7, 60e3), (7, 80e3), (7, 100e3)]
    wls = []
    for i in range(16):
        lev, rad = sizes[i % len(sizes)]
        wls.append(configs.random3d(300 + i, max_level=lev, n_freq=16, half_width=100e3,
                                    station_radius=rad, maxiter=args.maxiter))
    # Largest meshes first (finest level, then widest refinement radius):
    # the four largest start together and a worker that finishes takes the
    # next one, so the streams end on small ones.
    wls.sort(key=lambda wl: (wl.mesh.max_level, wl.mesh.station_refine_radius), reverse=True)
    n_streams = 4
    build_ctx = Context(0)
    contexts = [Context(0) for _ in range(n_streams)]
    t = time.time()
    jobs = []
    for wl in wls:
        z, err = multimodel.halfspace_data(wl)
        jobs.append(multimodel.ModelJob(wl, z, err))
    multimodel.prepare(jobs, [build_ctx], n_threads=16)
    t_build = time.time() - t
    edges = [j.model.n for j in jobs]
    log(f"[c3] 16 models built in {t_build:.1f}s, edges {edges}")
    prec = PREC_MIXED if args.c3_precision == "mixed" else 0
    prm = configs.ForwardParams(bicgstab_tol=1e-7, bicgstab_maxiter=args.maxiter).fast_params(
        precision=prec, dc_interval=args.dc_interval)
    # Device-resident gradients: the station steps between the forward and
    # the adjoint batch run on the device.
    wall = multimodel.run(jobs, n_streams, prm, prm, release=True, device=True, contexts=contexts)
    n_sys = multimodel.systems(jobs)
    iters = [int(sum(r["iterations"] for r in j.result.forward + j.result.adjoint)) for j in jobs]
    # A system is solved when it converged, or broke down (info 2) with its
    # true residual already below the tolerance.
    def solved(r):
        return bool(r["converged"]) or (r["info"] == 2 and r["true_residual"] < prm.tol)
    # Systems that stopped without meeting the tolerance: (model, phase,
    # system, info, true residual, iterations).
    unconverged = [(k, ph, i, r["info"], r["true_residual"], r["iterations"])
                   for k, j in enumerate(jobs)
                   for ph, rs in (("forward", j.result.forward), ("adjoint", j.result.adjoint))
                   for i, r in enumerate(rs) if not solved(r)]
    n_solved = n_sys - len(unconverged)
    conv = not unconverged
    out = {"metric": "solved systems/s, 16-model batch", "value": n_solved / wall, "unit": "systems/s",
           "systems": n_sys, "solved": n_solved, "models": 16, "frequencies": 16, "edges": edges, "total_edges": int(sum(edges)),
           "wall_s": wall, "build_s": t_build, "streams": n_streams, "converged": conv,
           "iterations_per_model": iters, "per_model_s": [j.seconds for j in jobs],
           "unconverged": unconverged,
           "unconverged_per_model": [{"forward": j.result.unconverged_forward,
                                      "adjoint": j.result.unconverged_adjoint} for j in jobs],
           "method": ("complex64 inner BiCGStab in FP64 refinement" if prec else "complex128 BiCGStab")
                     + " + SSOR with DivCorr every 100 iterations (forward and adjoint), device-resident "
                       "gradients",
           "dtype": "complex128 (f64)" if not prec else "complex64 inner, complex128 outer",
           "data": "synthetic (seeded random 3-D models, halfspace observations)"}
    log(f"[c3] {n_sys} systems in {wall:.1f}s, {n_solved} solved = {n_solved / wall:.1f} solved systems/s")
    for c in contexts + [build_ctx]:
        c.close()
    return out


# Layer shapes of the e0 network: (input channels, output channels) of the
# four 3x3x3 convolutions on a 32^3 grid.
NET_LAYERS = ((2, 32), (32, 32), (32, 32), (32, 12))


def network_leg(args, ctx):
    """BF16 MFMA throughput of the e0 network on a batch of 8 random samples:
    inference (forward) and one training step (forward, input and weight
    gradients, Adam at lr 0), timed over repeated blocking calls.  Useful
    flops per sample and layer 2 * 27 * Cin * Cout * 32^3; a training step
    adds the input gradient of layers 2-4 and the weight gradient of all
    four."""
    import torch

    from maple3dmt_amd.learned import CoarseNet
    dil = (1, 4, 8, 16)
    if os.path.exists(args.weights):
        with np.load(args.weights) as wts:
            dil = tuple(int(v) for v in wts["dilations"])
    batch, vox = 8, 32 ** 3
    net = CoarseNet(ctx, max_batch=batch, dilations=dil)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(1)
    d_in = torch.randn(batch * vox * 2, device=dev, dtype=torch.float32, generator=g)
    d_tgt = torch.randn(batch * vox * 12, device=dev, dtype=torch.float32, generator=g)
    d_out = torch.empty(batch * vox * 12, device=dev, dtype=torch.float32)
    torch.cuda.synchronize()
    per = 2 * 27 * vox
    f_fwd = batch * per * sum(ci * co for ci, co in NET_LAYERS)
    f_train = f_fwd + batch * per * (sum(ci * co for ci, co in NET_LAYERS[1:]) +
                                     sum(ci * co for ci, co in NET_LAYERS))
    reps = max(4, args.reps)
    for _ in range(3):
        net.forward(d_in.data_ptr(), d_out.data_ptr(), batch)
        net.train_step(d_in.data_ptr(), d_tgt.data_ptr(), batch, 0.0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        net.forward(d_in.data_ptr(), d_out.data_ptr(), batch)
    torch.cuda.synchronize()
    ms_fwd = (time.perf_counter() - t) * 1e3 / reps
    t = time.perf_counter()
    for _ in range(reps):
        net.train_step(d_in.data_ptr(), d_tgt.data_ptr(), batch, 0.0)
    torch.cuda.synchronize()
    ms_train = (time.perf_counter() - t) * 1e3 / reps
    net.close()
    tf_fwd = f_fwd / (ms_fwd * 1e-3) / 1e12
    tf_train = f_train / (ms_train * 1e-3) / 1e12
    out = {"batch": batch, "grid": 32, "layers": [list(x) for x in NET_LAYERS], "dilations": list(dil),
           "forward_ms": ms_fwd, "forward_gflop": f_fwd / 1e9, "forward_TFLOPs": tf_fwd,
           "train_step_ms": ms_train, "train_step_gflop": f_train / 1e9, "train_step_TFLOPs": tf_train,
           "peak_TFLOPs": MFMA_BF16_PEAK_TFLOPS, "forward_frac": tf_fwd / MFMA_BF16_PEAK_TFLOPS,
           "train_step_frac": tf_train / MFMA_BF16_PEAK_TFLOPS, "counters": None}
    try:
        with open(MFMA) as f:
            out["counters"] = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        pass
    log(f"[bench] network: forward {ms_fwd:.3f} ms ({tf_fwd:.1f} TFLOP/s), train step {ms_train:.3f} ms "
        f"({tf_train:.1f} TFLOP/s)")
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(model, wl, budget, gpu_iters):
    """The CPU oracle (the reference's forward solve: natural-order
    SSOR-BiCGStab from zero, dead-edge interpolation, post-hoc DivCorr;
    complex128) solving complete systems of the bench workload, highest
    frequency first, on one core until `budget` seconds are spent (at least
    the 100 Hz pair), then the same systems at once, one per core, each on
    its own oracle (the reference's frequency-parallel ranks).  value: that
    P-core throughput in solved systems per second.  The sample is the
    high-frequency end of the workload (19-100 Hz and below while the budget
    lasts); the low-frequency systems need thousands of natural-order
    iterations and do not fit a bounded run."""
    import threading

    from oracle.oracle import Oracle

    def make():
        o = Oracle(model, sigma_bg=wl.sigma_bg)
        o.set_sigma(model.array("sigma"))
        return o

    t = time.perf_counter()
    o = make()
    setup = time.perf_counter() - t
    sample, t1, its = [], [], []
    for f in sorted(wl.freqs, reverse=True):
        o.assemble(2 * np.pi * f)
        for pol in (0, 1):
            t = time.perf_counter()
            _, r = o.solve_pol(pol)
            t1.append(time.perf_counter() - t)
            its.append(int(r["iterations"]))
            sample.append((float(f), pol))
        if sum(t1) >= budget:
            break
    value_1 = len(sample) / sum(t1)
    P = len(sample)
    oracles = [None] * P

    def build(k):
        oracles[k] = make()
        oracles[k].assemble(2 * np.pi * sample[k][0])

    threads = [threading.Thread(target=build, args=(k,)) for k in range(P)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    start = threading.Barrier(P + 1)
    done = [0.0] * P

    def work(k):
        start.wait()
        _, rr = oracles[k].solve_pol(sample[k][1])
        done[k] = time.perf_counter()

    threads = [threading.Thread(target=work, args=(k,)) for k in range(P)]
    for th in threads:
        th.start()
    start.wait()
    t = time.perf_counter()
    for th in threads:
        th.join()
    wall = max(done) - t
    value_p = P / wall
    del oracles
    freqs = sorted({f for f, _ in sample}, reverse=True)
    fl = [float(v) for v in wl.freqs]
    gpu_its = [int(gpu_iters[2 * fl.index(f) + p]) for f, p in sample]
    desc = (f"complete forward solves (BiCGStab to 1e-7 + post-hoc DivCorr) of the {len(sample)} systems "
            f"at {', '.join(f'{f:.3g}' for f in freqs)} Hz: {sum(t1):.1f}s on one core "
            f"(natural-order iterations {its}; the GPU's multicolour SSOR needs {gpu_its} for them), {wall:.1f}s with one system per core on {P} cores; oracle setup {setup:.1f}s; "
            f"CPU {cpu_model()}")
    log(f"[bench] cpu baseline: 1 core {value_1:.4f} solves/s, {P} cores {value_p:.4f} solves/s; {desc}")
    return {"value": value_p, "unit": "solves/s", "cores": P, "kind": "port", "sample": desc,
            "single_core": {"value": value_1, "seconds": t1, "iterations": its},
            "gpu_multicolour_iterations": gpu_its,
            "systems": [{"freq_hz": f, "pol": p} for f, p in sample]}


if __name__ == "__main__":
    main()
