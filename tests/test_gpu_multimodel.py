"""Several models on one GPU, driven by concurrent host threads on their own
context streams, give exactly the gradients of running them one by one."""
import numpy as np
import pytest

from maple3dmt_amd import configs, multimodel
from maple3dmt_amd.native import Context

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
        # Solves stop as the reference's do: converged, or by its absolute
        # breakdown guards (|rho|, |omega| < 1e-30, bicgstab.cpp:79-85,
        # 152-157), which a tiny right-hand side can reach near the
        # tolerance, or (interleaved adjoint) by its stagnation rule
        # (forward_solver_fv.cpp:866-873); never by diverging.
        fw = [(r["converged"], r["info"], r["iterations"], r["true_residual"]) for r in a.result.forward]
        ad = [(r["converged"], r["diverged"], r["iterations"], r["true_residual"]) for r in a.result.adjoint]
        assert all(r["converged"] or r["true_residual"] < 1e-5 for r in a.result.forward), (a.workload.name, fw)
        assert not any(r["diverged"] for r in a.result.adjoint), (a.workload.name, ad)
        assert all(r["converged"] or r["true_residual"] < 1e-3 for r in a.result.adjoint), (a.workload.name, ad)
        assert [r["iterations"] for r in a.result.adjoint] == [r["iterations"] for r in b.result.adjoint]
        assert a.result.phi == b.result.phi
        assert np.array_equal(a.result.grad, b.result.grad)
        assert np.array_equal(a.result.z_pred, b.result.z_pred)
    for j in seq + par:
        j.model.close()
    for c in contexts:
        c.close()


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
