import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def solved(r, tol=1e-7):
    """A solve result that meets the tolerance: converged, or stopped by one
    of the reference's absolute breakdown guards (|rho|, |(r0, v)| or |(t, t)|
    below 1e-30, bicgstab.cpp:53, 79-85, 116-124) at an FP64 true residual
    already below tol.  Whether a trajectory meets a guard near convergence
    depends on rounding (reduction order, launch shape)."""
    return bool(r["converged"] or (r["info"] == 2 and r["true_residual"] < tol))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def ctx():
    from maple3dmt_amd import Context
    c = Context(0)
    yield c
    c.close()
