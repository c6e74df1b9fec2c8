"""Several models on one GPU, driven by concurrent host threads on their own
context streams, give exactly the gradients of running them one by one."""
import numpy as np
import pytest

from maple3dmt_amd import configs, multimodel
from maple3dmt_amd.native import Context
from conftest import solved
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _jobs():
    jobs = []
    for seed, level in ((21, 4), (22, 5), (23, 5)):
        wl = configs.random3d(seed, min_level=3, max_level=level, n_freq=2, half_width=50e3,
                              station_radius=5e3)
        z, err = multimodel.halfspace_data(wl)
        jobs.append(multimodel.ModelJob(wl, z, err))
    return jobs


def test_concurrent_models_match_sequential():
    contexts = [Context(0) for _ in range(3)]
    seq = _jobs()
    multimodel.prepare(seq, contexts[:1])
    multimodel.run(seq, 1)
    par = _jobs()
    multimodel.prepare(par, contexts)
    multimodel.run(par, 3)
    for a, b in zip(seq, par):
        # Forward solves meet the tolerance (converged, or stopped by one of
        # the reference's absolute breakdown guards at a true residual
        # already below it).  Interleaved adjoint solves end as the
        # reference's do: converged, or accepted by its stagnation rule
        # (a burst gaining less than 5 %, forward_solver_fv.cpp:866-873);
        # never diverged.
        fw = [(r["converged"], r["info"], r["iterations"], r["true_residual"]) for r in a.result.forward]
        ad = [(r["converged"], r["stagnated"], r["diverged"], r["true_residual"]) for r in a.result.adjoint]
        assert all(solved(r) for r in a.result.forward), (a.workload.name, fw)
        assert not any(r["diverged"] for r in a.result.adjoint), (a.workload.name, ad)
        assert all(solved(r) or r["stagnated"] for r in a.result.adjoint), (a.workload.name, ad)
        assert a.result.unconverged_forward == 0
        assert [r["iterations"] for r in a.result.adjoint] == [r["iterations"] for r in b.result.adjoint]
        assert a.result.phi == b.result.phi
        assert np.array_equal(a.result.grad, b.result.grad)
        assert np.array_equal(a.result.z_pred, b.result.z_pred)
    for j in seq + par:
        j.model.close()
    for c in contexts:
        c.close()


def test_dynamic_schedule_matches_sequential():
    """Models built on one context and taken from a shared queue by two
    workers on contexts of their own (each model moved to the stream of the
    worker that takes it) give exactly the gradients of running them one by
    one.  The build context is destroyed before the run: it stays alive
    until the last model has moved off it."""
    build = Context(0)
    workers = [Context(0) for _ in range(2)]
    seq = _jobs()
    multimodel.prepare(seq, workers[:1])
    multimodel.run(seq, 1)
    dyn = _jobs()
    multimodel.prepare(dyn, [build])
    build.close()
    multimodel.run(dyn, 2, contexts=workers)
    assert {j.worker for j in dyn} <= {0, 1}
    for a, b in zip(seq, dyn):
        assert b.model.ctx is workers[b.worker]
        assert [r["iterations"] for r in a.result.forward] == [r["iterations"] for r in b.result.forward]
        assert [r["iterations"] for r in a.result.adjoint] == [r["iterations"] for r in b.result.adjoint]
        assert a.result.phi == b.result.phi
        assert np.array_equal(a.result.grad, b.result.grad)
        assert np.array_equal(a.result.z_pred, b.result.z_pred)
    for j in seq + dyn:
        j.model.close()
    for c in workers:
        c.close()


def test_set_context_errors_and_same_context():
    """Moving an operator or station set: a null context is an argument
    error, a context already destroyed (kept alive by the objects still on
    it) a state error, the current context a no-op.  Moving the last object
    off a destroyed context frees it; the moved model then solves on its new
    stream."""
    from maple3dmt_amd.native import lib
    a, c = Context(0), Context(0)
    j1, j2 = _jobs()[:2]
    multimodel.prepare([j1], [a])
    multimodel.prepare([j2], [c])
    m1, m2 = j1.model, j2.model
    st = m1.device_stations(len(j1.workload.freqs))
    assert lib().m3d_operator_set_context(m1.device.handle, None) == 1      # M3D_ERR_ARG
    assert lib().m3d_stations_set_context(st.handle, None) == 1
    ch = c.handle
    c.close()                                   # destroyed; m2's operator keeps it alive
    assert lib().m3d_operator_set_context(m1.device.handle, ch) == 3        # M3D_ERR_STATE
    assert lib().m3d_stations_set_context(st.handle, ch) == 3
    m1.set_context(a)                           # its own context: nothing to do
    m2.set_context(a)                           # last object off c: c is freed
    assert m2.ctx is a
    for j in (j1, j2):
        E, Z, res = j.model.solve_frequencies(j.workload.freqs)
        assert all(solved(r) for r in res)
        j.model.close()
    a.close()


def test_release_after_gradient_keeps_results():
    """run(..., release=True) closes each model once its gradient is done
    (the device memory goes back) and leaves the gradients those of a run
    that keeps the models; the host threads build identical models."""
    contexts = [Context(0) for _ in range(2)]
    keep = _jobs()
    multimodel.prepare(keep, contexts, n_threads=1)
    multimodel.run(keep, 2)
    rel = _jobs()
    multimodel.prepare(rel, contexts, n_threads=3)
    multimodel.run(rel, 2, release=True)
    for a, b in zip(keep, rel):
        assert b.model.handle is None
        assert a.result.phi == b.result.phi
        assert np.array_equal(a.result.grad, b.result.grad)
    for j in keep:
        j.model.close()
    for c in contexts:
        c.close()


def test_trim_frees_workspaces_and_keeps_results():
    """m3d_operator_trim frees the solve workspaces (the device bytes they
    held go back to zero) and the next solve, which allocates them again,
    gives the same fields bit for bit."""
    ctx = Context(0)
    wl = configs.random3d(24, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    b0 = m.device.device_bytes()
    E1, Z1, r1 = m.solve_frequencies(wl.freqs)
    b1 = m.device.device_bytes()
    print("device bytes", b0, b1, flush=True)
    assert b1["operator"] == b0["operator"] > 0
    assert b1["workspaces"] > 0
    m.device.trim()
    assert m.device.device_bytes() == {"operator": b0["operator"], "workspaces": 0}
    E2, Z2, r2 = m.solve_frequencies(wl.freqs)
    assert [r["iterations"] for r in r1] == [r["iterations"] for r in r2]
    assert np.array_equal(E1, E2)
    assert np.array_equal(Z1, Z2)
    m.close()
    ctx.close()


def test_uneven_models_concurrent_gradients_match_oracle():
    """The 16-model batch in small: three models on meshes of uneven size
    (levels 3-4 and 3-5, two station radii), each with its forward and
    adjoint systems, run concurrently on three streams with the station
    steps on the device (multimodel.run(device=True)); each model's misfit
    and gradient against the oracle's per-frequency loop
    (inversion_fv.cpp:250-290), to the bar of the single-model gradient
    test (2e-3)."""
    contexts = [Context(0) for _ in range(3)]
    jobs, oracles = [], []
    for seed, level, radius in ((31, 4, 5e3), (32, 5, 5e3), (33, 5, 12e3)):
        wl = configs.random3d(seed, min_level=3, max_level=level, n_freq=2, half_width=50e3,
                              station_radius=radius)
        jobs.append(multimodel.ModelJob(wl, None, None))
    multimodel.prepare(jobs, contexts)
    sizes = [j.model.n for j in jobs]
    print("edges per model", sizes, flush=True)
    assert len(set(sizes)) == 3
    for j in jobs:
        m, wl = j.model, j.workload
        o = Oracle(m, sigma_bg=wl.sigma_bg)
        o.set_sigma(m.array("sigma"))
        o.find_stations(wl.stations[:, :2])
        zs = [o.solve_frequency(f)[2] for f in wl.freqs]
        j.z_obs = np.stack(zs) * (1.05 + 0.02j)
        floor = 0.05 * np.maximum(np.abs(j.z_obs[:, :, 1]), np.abs(j.z_obs[:, :, 2]))
        j.z_err = np.repeat(floor[:, :, None], 4, axis=2)
        oracles.append(o)
    multimodel.run(jobs, 3, device=True)
    for j, o in zip(jobs, oracles):
        res = j.result
        assert all(solved(r) for r in res.forward), res.forward
        assert all(solved(r) or r["stagnated"] for r in res.adjoint), res.adjoint
        phi_o, g_o = 0.0, np.zeros(j.model.nc)
        for fi, f in enumerate(j.workload.freqs):
            e1, e2, z, _ = o.solve_frequency(f)
            r = j.z_obs[fi] - z
            phi_o += 0.5 * float(np.sum(np.abs(r) ** 2 / j.z_err[fi] ** 2))
            q1, q2 = o.adjoint_rhs(np.conj(r) / j.z_err[fi] ** 2)
            g_o += o.sensitivity(e1, o.adjoint(q1)[0]) + o.sensitivity(e2, o.adjoint(q2)[0])
        print(j.workload.name, "phi", res.phi, phi_o, "grad rel err",
              np.linalg.norm(res.grad - g_o) / np.linalg.norm(g_o), flush=True)
        assert abs(res.phi - phi_o) <= 2e-3 * phi_o
        assert np.linalg.norm(res.grad - g_o) <= 2e-3 * np.linalg.norm(g_o)
    for j in jobs:
        j.model.close()
    for c in contexts:
        c.close()
