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
        assert all(r["converged"] for r in a.result.forward + a.result.adjoint)
        assert a.result.phi == b.result.phi
        assert np.array_equal(a.result.grad, b.result.grad)
        assert np.array_equal(a.result.z_pred, b.result.z_pred)
    for j in seq + par:
        j.model.close()
    for c in contexts:
        c.close()
