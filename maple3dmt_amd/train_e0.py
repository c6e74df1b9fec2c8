"""Train the coarse-grid e0 network on solver-generated solutions and
evaluate it against the zero initial guess.

    python -m maple3dmt_amd.train_e0 --train-models 200 --eval-models 20 \
        --steps 1500 --out maple3dmt_amd/data/e0_net.npz

Training models are seeded random 3-D models (three layers plus blocks) on
their own octree meshes; each contributes its 8 frequencies x 2
polarisations.  Evaluation solves held-out models twice, from zero and from
the predicted e0 (zero-guess fallback where ||b - A e0|| >= ||b||), at the
same tolerance, and reports iterations and solves/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import configs
from .learned import Box, CoarseNet, initial_guess, make_sample, train, training_workload
from .native import MODE_FORWARD, Context

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def evaluate(ctx, net, box, seeds, max_level, repeats=1):
    import torch
    dev = torch.device("cuda:0")
    rows = []
    for seed in seeds:
        wl = configs.random3d(seed, max_level=max_level, half_width=100e3, station_radius=20e3)
        m = configs.build_model(wl, ctx)
        omegas = np.repeat(2 * np.pi * wl.freqs, 2)
        ns = len(omegas)
        rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
        aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(omegas)])
        d_rhs = torch.from_numpy(rhs).to(dev)
        d_aux = torch.from_numpy(aux).to(dev)
        d_x = torch.zeros_like(d_rhs)
        prm0 = wl.fwd.solve_params()
        prm1 = wl.fwd.solve_params()
        prm1.use_initial_guess = 2
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            r0 = m.device.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(),
                                       d_x.data_ptr(), ns, prm0)
        torch.cuda.synchronize()
        dt0 = (time.perf_counter() - t0) / repeats
        x_zero = d_x.cpu().numpy()
        t1 = time.perf_counter()
        for _ in range(repeats):
            x0 = initial_guess(net, m, wl, box)       # network + prolongation on the device
            r1 = m.device.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(),
                                       x0.data_ptr(), ns, prm1)
        torch.cuda.synchronize()
        dt1 = (time.perf_counter() - t1) / repeats
        x_learn = x0.cpu().numpy()
        dz = []
        for f in range(len(wl.freqs)):
            za = m.extract_impedance(omegas[2 * f], x_zero[2 * f], x_zero[2 * f + 1])
            zb = m.extract_impedance(omegas[2 * f], x_learn[2 * f], x_learn[2 * f + 1])
            dz.append(float(np.max(np.abs(zb[:, 1:3] - za[:, 1:3]) / np.abs(za[:, 1:3]))))
        row = dict(seed=seed, edges=m.n,
                   iters_zero=[r["iterations"] for r in r0],
                   iters_e0=[r["iterations"] for r in r1],
                   guess_used=[r["guess_used"] for r in r1],
                   converged=all(r["converged"] for r in r0 + r1),
                   s_zero=dt0, s_e0=dt1, max_rel_dZ=max(dz))
        log(f"[eval] seed {seed}: {m.n} edges, iterations {sum(row['iters_zero'])} -> "
            f"{sum(row['iters_e0'])}, {dt0:.2f}s -> {dt1:.2f}s, dZ {row['max_rel_dZ']:.1e}")
        rows.append(row)
        m.close()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-models", type=int, default=200)
    ap.add_argument("--eval-models", type=int, default=20)
    ap.add_argument("--train-level", type=int, default=5)
    ap.add_argument("--eval-level", type=int, default=6)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(DATA, "e0_net.npz"))
    ap.add_argument("--weights", default=None, help="skip training, load these weights")
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)

    import torch  # noqa: F401  (device memory for training batches)
    ctx = Context(0)
    box = Box()
    net = CoarseNet(ctx, max_batch=max(a.batch, 8), seed=1)
    report = {"box": box.arr().tolist(), "train_models": a.train_models, "train_level": a.train_level,
              "eval_level": a.eval_level}
    if a.weights:
        d = np.load(a.weights)
        net.set_params(d["w"], d["b"])
    else:
        t0 = time.time()
        samples = []
        gen_solve_iters = 0
        for k in range(a.train_models):
            wl = training_workload(10_000 + k, max_level=a.train_level)
            m = configs.build_model(wl, ctx)
            s = make_sample(m, wl, box)
            gen_solve_iters += sum(s.iterations)
            samples.append(s)
            m.close()
            if k % 20 == 0:
                log(f"[data] model {k}: {m.n} edges, {time.time() - t0:.0f}s")
        report["data_seconds"] = time.time() - t0
        report["data_solver_iterations"] = gen_solve_iters
        t1 = time.time()
        losses = train(net, samples, a.steps, a.lr, a.batch, log=log)
        report["train_seconds"] = time.time() - t1
        report["loss_first"] = float(np.mean(losses[:20]))
        report["loss_last"] = float(np.mean(losses[-20:]))
        w, b = net.params()
        os.makedirs(os.path.dirname(a.out), exist_ok=True)
        np.savez(a.out, w=w, b=b, box=box.arr(), dilations=np.array(net.dil))
        log(f"[train] loss {report['loss_first']:.3e} -> {report['loss_last']:.3e}; saved {a.out}")
    rows = evaluate(ctx, net, box, [50_000 + k for k in range(a.eval_models)], a.eval_level)
    it0 = sum(sum(r["iters_zero"]) for r in rows)
    it1 = sum(sum(r["iters_e0"]) for r in rows)
    s0 = sum(r["s_zero"] for r in rows)
    s1 = sum(r["s_e0"] for r in rows)
    report["eval"] = rows
    report["summary"] = {"iterations_zero": it0, "iterations_e0": it1,
                         "iteration_ratio": it1 / max(it0, 1), "seconds_zero": s0, "seconds_e0": s1,
                         "speedup": s0 / max(s1, 1e-9),
                         "max_rel_dZ": max(r["max_rel_dZ"] for r in rows),
                         "all_converged": all(r["converged"] for r in rows)}
    log(f"[eval] {json.dumps(report['summary'])}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(report["summary"]), flush=True)


if __name__ == "__main__":
    main()
