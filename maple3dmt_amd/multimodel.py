"""Misfit gradients of several models (meshes of uneven size) on one GPU.

Each model keeps its own device operator resident and runs the gradient's two
batched phases (every forward system, then every adjoint system; see
inversion.py) on the stream of its own context.  A pool of host threads
drives the models, so the GPU overlaps one model's narrow low-frequency tail
with the wide phases of others.  ctypes releases the GIL inside every library
call, and every context, operator and workspace belongs to one thread at a
time, so the threads share nothing but the GPU.

This replaces the reference's MPI frequency parallelism
(src/utils/freq_parallel.cpp) for the many-model workload: there, one rank per
frequency runs the per-frequency loop of InversionFV::compute_gradient_
(src/inversion/inversion_fv.cpp:230-312) on the CPU.
"""
from __future__ import annotations

import os
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import configs, inversion
from .native import Context, Model, SolveParams


@dataclass
class ModelJob:
    """One model of the batch: its workload, observed data and errors."""
    workload: configs.Workload
    z_obs: np.ndarray                  # [nf, stations, 4]
    z_err: np.ndarray                  # [nf, stations, 4]
    model: Optional[Model] = None      # built by prepare()
    result: Optional[inversion.GradientResult] = None
    seconds: float = 0.0
    worker: int = -1


def build_models(workloads: Sequence[configs.Workload], contexts: Sequence[Optional[Context]],
                 n_threads: Optional[int] = None) -> List:
    """Models of several workloads built at once, one host thread each (the
    host builder holds no shared state and its calls release the GIL);
    workload i goes to context i % len(contexts) (None: host-only).  The
    models are identical to ones built one after another."""
    n = len(workloads)
    if n == 0:
        return []
    n_threads = n_threads or min(n, os.cpu_count() or 1)
    out: List = [None] * n
    with ThreadPoolExecutor(max_workers=n_threads) as ex:
        futs = {ex.submit(configs.build_model, wl, contexts[i % len(contexts)]): i
                for i, wl in enumerate(workloads)}
        for f, i in futs.items():
            out[i] = f.result()
    return out


def prepare(jobs: Sequence[ModelJob], contexts: Sequence[Context], n_threads: Optional[int] = None) -> None:
    """Build every model's mesh, operators and device operator (host work,
    outside any timed region), several at once; job i is assigned to context
    i % len(contexts)."""
    for i, job in enumerate(jobs):
        job.worker = i % len(contexts)
    models = build_models([j.workload for j in jobs], contexts, n_threads)
    for job, m in zip(jobs, models):
        job.model = m


def run(jobs: Sequence[ModelJob], n_workers: int, params: Optional[SolveParams] = None,
        adjoint_params: Optional[SolveParams] = None, release: bool = False,
        device: bool = False, contexts: Optional[Sequence[Context]] = None) -> float:
    """Gradient of every job; with one worker all jobs run in order on the
    calling thread, else worker w runs, in order, the jobs whose context
    index is w modulo n_workers (so every job runs, and a context's jobs stay
    on one thread).  With contexts (one per worker; n_workers must equal
    their number) the workers instead take the jobs in list order from one
    queue, and each job's model moves to the taking worker's context first
    (Model.set_context), so a worker that finishes early takes the next job
    whatever context the model was built on; build the models on a context
    no worker uses, so that the move does not wait for another worker's
    solves.  device: the device-resident gradient (inversion.gradient_device)
    instead of the host-station one.  With release, each model (and its
    device memory) is closed once its gradient is done.  Returns the wall
    time in seconds."""
    if n_workers < 1:
        raise ValueError("n_workers must be at least 1")
    if contexts is not None and len(contexts) != n_workers:
        raise ValueError("dynamic scheduling takes one context per worker")
    if any(job.worker < 0 or job.model is None for job in jobs):
        raise ValueError("every job must be prepared (prepare()) before run()")
    errors: List[BaseException] = []
    grad = inversion.gradient_device if device else inversion.gradient
    pending: "queue.Queue[ModelJob]" = queue.Queue()
    if contexts is not None:
        for job in jobs:
            pending.put(job)

    def solve(job: ModelJob):
        t = time.perf_counter()
        job.result = grad(job.model, job.workload.freqs, job.z_obs, job.z_err, params, adjoint_params)
        job.seconds = time.perf_counter() - t
        if release:
            job.model.close()

    def work(w: int):
        try:
            if contexts is not None:
                while not errors:
                    try:
                        job = pending.get_nowait()
                    except queue.Empty:
                        return
                    job.worker = w
                    job.model.set_context(contexts[w])
                    solve(job)
                return
            for job in jobs:
                if job.worker % n_workers == w:
                    solve(job)
        except BaseException as e:  # re-raised in the caller
            errors.append(e)

    t0 = time.perf_counter()
    threads = [threading.Thread(target=work, args=(w,)) for w in range(n_workers)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return wall


def systems(jobs: Sequence[ModelJob]) -> int:
    """Solved systems: two forward and two adjoint per frequency and model."""
    return sum(4 * len(job.workload.freqs) for job in jobs)


def halfspace_data(workload: configs.Workload, rel_err: float = 0.05):
    """Synthetic observations for benchmarking: the impedance of the
    workload's background halfspace at every station, Zxy = -Zyx =
    sqrt(i w mu0 / sigma_bg) (halfspace_1d_test.cpp:124-125), with errors
    rel_err |Z| on every component."""
    mu0 = 4e-7 * np.pi
    ns = len(workload.stations)
    nf = len(workload.freqs)
    z = np.zeros((nf, ns, 4), np.complex128)
    for f, fr in enumerate(workload.freqs):
        zz = np.sqrt(1j * 2 * np.pi * fr * mu0 / workload.sigma_bg)
        z[f, :, 1] = zz
        z[f, :, 2] = -zz
    err = np.repeat(rel_err * np.abs(z[:, :, 1:2]), 4, axis=2)
    return z, err
