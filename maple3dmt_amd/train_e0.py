"""Train the coarse-grid e0 network on solver-generated solutions, and
evaluate what it gains over the zero initial guess.

    python -m maple3dmt_amd.train_e0 train --models 2000 --levels 5,6,7 --box domain \
        --steps 20000 --out maple3dmt_amd/data/e0_net.npz
    python -m maple3dmt_amd.train_e0 eval --weights maple3dmt_amd/data/e0_net.npz \
        --models 200 --levels 5,6,7 --report e0_eval.json
    python -m maple3dmt_amd.train_e0 bound --seed 1 --level 7 --report e0_bound.json

Training models are seeded random 3-D models (three layers plus one to four
blocks, configs.random3d) on their own octree meshes, coarsest level 3 and
finest level 5, 6 or 7 drawn per seed; each contributes its 8 frequencies.
The voxel box covers the whole mesh domain (--box domain) or the 80 km box
around the stations (--box station).  Targets are the solver's own E_sec = E - E_pri (mixed-precision
solve with interleaved DivCorr to a relative residual of 1e-6; models with an
unconverged system are skipped), deposited on the network's voxel grid and
scaled to unit RMS per polarisation: the network predicts the shape, the
solve's least-squares scaling of the guess restores the amplitude.

Evaluation solves held-out models (seeds disjoint from training) at the same
tolerance: the plain FP64 method from zero (the reference), and the fast path
(mixed precision + interleaved DivCorr) from zero and from the network's e0.
The e0 start is the least-squares scaled guess (use_initial_guess = 2), kept
per system only where it at least halves the residual, else the zero start.
Reported per model: BiCGStab iterations per system, solve wall time (the e0
variant includes the network inference and prolongation), the scaled guess's
residual and the largest |dZ|/|Z| against the reference solve.

`bound` measures what any coarse-grid e0 can gain: the converged solution
itself, restricted to the voxel grid and prolonged back, as the start.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import configs
from .learned import (Box, CoarseNet, domain_box, initial_guess, sample_from_fields, seeded_level, solve_model,
                      train)
from .native import MODE_FORWARD, PREC_MIXED, Context

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
TRAIN_SEED0 = 10_000
EVAL_SEED0 = 50_000


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _levels(a):
    return tuple(int(v) for v in a.levels.split(",")), tuple(float(v) for v in a.level_weights.split(","))


def _workload(seed, a, maxiter=20000):
    levels, weights = _levels(a)
    lev = seeded_level(seed, levels, weights)
    return configs.random3d(seed, max_level=lev, half_width=100e3, station_radius=20e3, maxiter=maxiter)


def cmd_train(a):
    """Data generation and training in one process.  Models are built on
    host threads one batch ahead of the solves, alternating between two
    contexts so that no build shares a stream with a running solve; each
    model's fields go to a host thread for the voxel deposition while the next
    model solves.  Generation stops after --models models or --data-seconds,
    whichever comes first."""
    import torch  # noqa: F401  (device memory for the training batches)
    from concurrent.futures import ThreadPoolExecutor
    ctxs = [Context(0), Context(0)]
    dil = tuple(int(v) for v in a.dilations.split(","))
    net = CoarseNet(ctxs[0], max_batch=max(a.batch, 8), dilations=dil, seed=a.seed)
    t0 = time.time()
    pool = ThreadPoolExecutor(max_workers=a.threads)      # voxel deposition
    builders = ThreadPoolExecutor(max_workers=a.build_batch)
    batcher = ThreadPoolExecutor(max_workers=1)
    B = a.build_batch
    seeds = [TRAIN_SEED0 + k for k in range(a.models)]

    def build_batch(i):
        wls = [_workload(sd, a) for sd in seeds[i:i + B]]
        ctx = ctxs[(i // B) % 2]
        return list(zip(wls, builders.map(lambda w: configs.build_model(w, ctx), wls)))

    samples, its, edges, levels_used, skipped, pending = [], 0, [], [], 0, []
    box0 = None
    nxt = batcher.submit(build_batch, 0)
    i = 0
    while i < len(seeds):
        batch = nxt.result()
        i += B
        over = time.time() - t0 > a.data_seconds
        nxt = batcher.submit(build_batch, i) if i < len(seeds) and not over else None
        for wl, m in batch:
            box = domain_box(wl) if a.box == "domain" else Box()
            box0 = box0 or box
            esec, res = solve_model(m, wl)
            if not all(r["converged"] for r in res):
                skipped += 1
                m.close()
                continue
            its += sum(r["iterations"] for r in res)
            edges.append(m.n)
            levels_used.append(wl.mesh.max_level)
            pending.append((pool.submit(sample_from_fields, m, wl, box, esec, res), m))
        still = []
        for f, m in pending:   # one pass: a future is either collected or kept
            if f.done():
                samples.append(f.result())
                m.close()
            else:
                still.append((f, m))
        pending = still
        log(f"[data] {len(samples) + len(pending)} models ({skipped} skipped), {sum(edges)} edges, "
            f"{time.time() - t0:.0f}s")
        if nxt is None:
            break
    for f, m in pending:
        samples.append(f.result())
        m.close()
    for ex in (pool, builders, batcher):
        ex.shutdown()
    t_data = time.time() - t0
    log(f"[data] {len(samples)} models x 8 frequencies in {t_data:.0f}s; levels "
        f"{ {l: levels_used.count(l) for l in sorted(set(levels_used))} }")
    t1 = time.time()
    losses = train(net, samples, a.steps, a.lr, a.batch, seed=a.seed, log=log)
    t_train = time.time() - t1
    w, b = net.params()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.savez(a.out, w=w, b=b, box=box0.arr(), dilations=np.array(net.dil), box_kind=np.array(a.box))
    rep = {"models": len(samples), "models_requested": a.models, "skipped_unconverged": skipped,
           "levels": {str(l): levels_used.count(l) for l in sorted(set(levels_used))},
           "edges_mean": float(np.mean(edges)), "edges_max": int(np.max(edges)),
           "samples": len(samples) * 8, "box": a.box, "data_seconds": t_data,
           "data_solver_iterations": its, "steps": a.steps, "batch": a.batch, "lr": a.lr,
           "train_seconds": t_train, "loss_first": float(np.mean(losses[:20])),
           "loss_last": float(np.mean(losses[-20:]))}
    log(f"[train] {json.dumps(rep)}; saved {a.out}")
    print(json.dumps(rep), flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    net.close()
    for c in ctxs:
        c.close()


def _solve(m, omegas, d_rhs, d_aux, d_x, prm):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = m.device.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(), d_x.data_ptr(),
                              len(omegas), prm)
    torch.cuda.synchronize()
    return r, time.perf_counter() - t


def cmd_eval(a):
    import torch
    dev = torch.device("cuda:0")
    ctx = Context(0)
    d = np.load(a.weights)
    box = Box(*d["box"].tolist())
    net = CoarseNet(ctx, max_batch=8, dilations=tuple(int(v) for v in d["dilations"]))
    net.set_params(d["w"], d["b"])
    rows = []
    t_start = time.time()
    for k in range(a.models):
        if time.time() - t_start > a.seconds:
            log(f"[eval] time budget reached after {k} models")
            break
        seed = EVAL_SEED0 + k
        wl = _workload(seed, a, a.maxiter)
        m = configs.build_model(wl, ctx)
        nf = len(wl.freqs)
        omegas = np.repeat(2 * np.pi * wl.freqs, 2)
        rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
        aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(omegas)])
        d_rhs = torch.from_numpy(rhs).to(dev)
        d_aux = torch.from_numpy(aux).to(dev)
        # Untimed warm-up: the first solve on a new model allocates its
        # workspaces and captures its graphs; without it the first timed
        # start would pay for them.
        _solve(m, omegas, d_rhs, d_aux, torch.zeros_like(d_rhs), wl.fwd.fast_params())
        row = {"seed": seed, "edges": m.n, "level": wl.mesh.max_level}
        zref = None
        legs = (("fast", PREC_MIXED, ("zero", "e0")),) if a.no_fp64 else \
            (("fp64", 0, ("zero",)), ("fast", PREC_MIXED, ("zero", "e0")))
        for prec_name, prec, starts in legs:
            for start in starts:
                prm = wl.fwd.fast_params() if prec else wl.fwd.solve_params()
                torch.cuda.synchronize()
                t = time.perf_counter()
                if start == "zero":
                    d_x = torch.zeros_like(d_rhs)
                    r, dt = _solve(m, omegas, d_rhs, d_aux, d_x, prm)
                    dt_net = 0.0
                else:
                    d_x = initial_guess(net, m, wl, box)
                    torch.cuda.synchronize()
                    dt_net = time.perf_counter() - t
                    prm.use_initial_guess = 2
                    r, dt = _solve(m, omegas, d_rhs, d_aux, d_x, prm)
                x = d_x.cpu().numpy()
                z = np.stack([m.extract_impedance(omegas[2 * f], x[2 * f], x[2 * f + 1]) for f in range(nf)])
                if zref is None:
                    zref = z
                off = (slice(None), slice(None), slice(1, 3))
                key = f"{prec_name}_{start}"
                row[key] = {"iterations": [q["iterations"] for q in r],
                            "guess_used": [q["guess_used"] for q in r],
                            "guess_residual": [q["guess_residual"] for q in r],
                            "converged": all(q["converged"] for q in r),
                            "seconds": dt + dt_net, "net_seconds": dt_net,
                            "max_rel_dZ": float(np.max(np.abs(z[off] - zref[off]) / np.abs(zref[off])))}
        log(f"[eval] seed {seed}: level {wl.mesh.max_level}, {m.n} edges; iterations fast "
            f"{sum(row['fast_zero']['iterations'])} -> e0 {sum(row['fast_e0']['iterations'])}; seconds fast "
            f"{row['fast_zero']['seconds']:.2f} -> e0 {row['fast_e0']['seconds']:.2f}; guess used "
            f"{sum(row['fast_e0']['guess_used'])}/16, residual {np.round(row['fast_e0']['guess_residual'], 3).tolist()}")
        rows.append(row)
        m.close()
    keys = ("fast_zero", "fast_e0") if a.no_fp64 else ("fp64_zero", "fast_zero", "fast_e0")

    def summarise(rs):
        out = {}
        for key in keys:
            out[key] = {"iterations": int(sum(sum(r[key]["iterations"]) for r in rs)),
                        "seconds": float(sum(r[key]["seconds"] for r in rs)),
                        "solves_per_s": float(16 * len(rs) / sum(r[key]["seconds"] for r in rs)),
                        "all_converged": all(r[key]["converged"] for r in rs),
                        "guess_used_fraction": float(np.mean([np.mean(r[key]["guess_used"]) for r in rs])),
                        "max_rel_dZ": float(max(r[key]["max_rel_dZ"] for r in rs))}
        return out

    summ = summarise(rows)
    by_level = {str(l): summarise([r for r in rows if r["level"] == l]) for l in sorted({r["level"] for r in rows})}
    rep = {"weights": os.path.basename(a.weights), "models": len(rows), "levels": a.levels,
           "summary": summ, "by_level": by_level, "per_model": rows}
    log(f"[eval] summary {json.dumps(summ)}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    print(json.dumps(summ), flush=True)
    net.close()
    ctx.close()


def cmd_bound(a):
    """Upper bound of any coarse-grid e0: the converged E_sec itself,
    deposited on the 32^3 voxel grid (the network's output space) and prolonged
    back to the edges, used as the start of the plain FP64 method and of the
    fast path (unscaled and least-squares scaled).  A trained network cannot
    start closer to the solution than this restriction of the solution."""
    import torch
    from .learned import Box, CoarseNet, deposit
    dev = torch.device("cuda:0")
    wl = configs.random3d(a.seed, max_level=a.level, half_width=a.half_width,
                          station_radius=a.station_radius, maxiter=a.maxiter)
    ctx = Context(0)
    m = configs.build_model(wl, ctx)
    nf = len(wl.freqs)
    omegas = np.repeat(2 * np.pi * wl.freqs, 2)
    rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
    aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(omegas)])
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_aux = torch.from_numpy(aux).to(dev)
    report = {"workload": wl.name, "edges": m.n, "starts": {}}
    x = torch.zeros_like(d_rhs)
    r, dt = _solve(m, omegas, d_rhs, d_aux, x, wl.fwd.fast_params())
    esec = x.cpu().numpy() - aux
    net = CoarseNet(ctx, max_batch=nf)
    exyz = torch.from_numpy(np.stack([m.array("edge_x"), m.array("edge_y"), m.array("edge_z")],
                                     1).astype(np.float32)).to(dev)
    edir = torch.from_numpy(m.array("edge_dir")).to(dev)
    boxes = {"station_box": Box(), "whole_domain": Box(*list(m.mesh_params.domain))}
    for bname, box in boxes.items():
        dep = np.stack([deposit(m, box, esec[2 * f], esec[2 * f + 1]) for f in range(nf)])
        d_dep = torch.from_numpy(dep.astype(np.float32)).to(dev)
        x0 = torch.empty((2 * nf, m.n), dtype=torch.complex128, device=dev)
        net.prolong(d_dep.data_ptr(), nf, m.n, exyz.data_ptr(), edir.data_ptr(), box, x0.data_ptr())
        torch.cuda.synchronize()
        g = x0.cpu().numpy()
        entry = {"rel_l2_error": [round(float(np.linalg.norm(g[i] - esec[i]) / np.linalg.norm(esec[i])), 4)
                                  for i in range(2 * nf)]}
        for method, mk in (("fp64", wl.fwd.solve_params), ("fast", wl.fwd.fast_params)):
            for start in (0, 1, 3):
                prm = mk()
                prm.use_initial_guess = start
                xx = x0.clone() if start else torch.zeros_like(d_rhs)
                r, dt = _solve(m, omegas, d_rhs, d_aux, xx, prm)
                entry[f"{method}_mode{start}"] = {
                    "seconds": dt, "iterations": [q["iterations"] for q in r],
                    "guess_used": [int(q["guess_used"]) for q in r],
                    "guess_residual": [round(q["guess_residual"], 4) for q in r]}
                print(bname, method, start, round(dt, 3), entry[f"{method}_mode{start}"]["iterations"], flush=True)
        report["starts"][bname] = entry
    net.close()
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--models", type=int, default=2000)
    t.add_argument("--levels", default="5,6,7", help="finest octree levels of the models (coarsest 3)")
    t.add_argument("--level-weights", default="0.5,0.35,0.15")
    t.add_argument("--box", choices=("domain", "station"), default="domain")
    t.add_argument("--data-seconds", type=float, default=1e9, help="stop generating data after this long")
    t.add_argument("--threads", type=int, default=12)
    t.add_argument("--build-batch", type=int, default=12)
    t.add_argument("--report", default=None)
    t.add_argument("--steps", type=int, default=6000)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=5e-4)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--dilations", default="1,2,4,1", help="dilation of each of the four conv layers")
    t.add_argument("--out", default=os.path.join(DATA, "e0_net.npz"))
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=os.path.join(DATA, "e0_net.npz"))
    e.add_argument("--models", type=int, default=200)
    e.add_argument("--levels", default="5,6,7")
    e.add_argument("--level-weights", default="0.5,0.35,0.15")
    e.add_argument("--seconds", type=float, default=1e9, help="stop after this long")
    e.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 reference solve")
    e.add_argument("--maxiter", type=int, default=40000)
    e.add_argument("--report", default=None)
    bd = sub.add_parser("bound")
    bd.add_argument("--seed", type=int, default=1)
    bd.add_argument("--level", type=int, default=7)
    bd.add_argument("--half-width", type=float, default=100e3)
    bd.add_argument("--station-radius", type=float, default=80e3)
    bd.add_argument("--maxiter", type=int, default=40000)
    bd.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    {"train": cmd_train, "eval": cmd_eval, "bound": cmd_bound}[a.cmd](a)


if __name__ == "__main__":
    main()
