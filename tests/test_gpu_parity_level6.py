"""Forward impedance parity on level-6 random models (octree levels 3-6,
39-66 K edges, 121 stations, 8 frequencies 1e-3..1e2 Hz, both
polarisations) for every solve path: plain FP64 (the reference's method),
complex64 inner solves in FP64 refinement, FP64 with interleaved DivCorr, and
the fast path (both).  The oracle's impedances come from committed fixtures
(tests/golden/level6_seed*.npz, written by make_level6_fixtures.py with the
CPU oracle at the reference's tolerance 1e-7 and converged at 1e-10).

What the reference's tolerance fixes.  The relative residual 1e-7 of the
reference (forward_solver_fv.cpp:347-354) leaves the oracle's own Z up to
1.0e-4 away from its converged value on these meshes at 19-100 Hz, where the
skin depth (~0.5 km) is far below the 3 km station cells and BiCGStab meets
1e-7 in two to ten iterations.  Any solver that does not replay the
oracle's rounding lands elsewhere inside that spread, so the bar is stated
against what the tolerance determines:

  * converged (device tol 1e-9): |dZ|/|Z| <= 1e-5 against the converged
    oracle, rho_a within 0.1 %, phase within 0.05 degrees, per frequency and
    station, off-diagonal Z (the diagonal to 1e-5 of the largest |Z|);
  * at the reference's tol 1e-7: |dZ|/|Z| against the oracle at 1e-7 within
    max(1e-5, 3 x the oracle's own spread at that frequency), where the
    spread is max |Z_oracle(1e-7) - Z_oracle(1e-10)| / |Z_oracle(1e-10)|.

Each case prints its largest error per frequency.
"""
import os

import numpy as np
import pytest

from maple3dmt_amd import configs
from maple3dmt_amd.native import PREC_MIXED
from conftest import solved

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = (50001, 50003)
PATHS = ("fp64", "mixed", "interleaved", "fast")
MU0 = 4e-7 * np.pi


@pytest.fixture(scope="module", params=SEEDS)
def level6(request, ctx):
    seed = request.param
    fx = np.load(os.path.join(HERE, "golden", f"level6_seed{seed}.npz"))
    wl = configs.random3d(seed, max_level=6, half_width=100e3, station_radius=20e3, maxiter=40000)
    m = configs.build_model(wl, ctx)
    fp = np.array([m.n, m.nc, m.nn, m.array("edge_x").sum(), m.array("edge_y").sum(),
                   m.array("edge_z").sum()])
    assert np.allclose(fp, fx["fingerprint"], rtol=0, atol=1e-6), "mesh differs from the fixture's"
    assert np.allclose(fx["freqs"], wl.freqs)
    yield seed, wl, m, fx["Z"]
    m.close()


def _params(m, path, tol):
    if path == "fp64":
        prm = m.fwd_params.solve_params(tol=tol)
    elif path == "mixed":
        prm = m.fwd_params.solve_params(precision=PREC_MIXED, inner_tol=1e-3, tol=tol)
    elif path == "interleaved":
        prm = m.fwd_params.solve_params(tol=tol)
        prm.interleaved_forward = 1
    else:
        prm = m.fwd_params.fast_params(tol=tol)
    prm.maxiter = 40000
    return prm


def _rel(z, zr):
    return np.abs(z[..., 1:3] - zr[..., 1:3]) / np.abs(zr[..., 1:3])


@pytest.mark.parametrize("path", PATHS)
def test_level6_converged_parity(level6, path):
    seed, wl, m, Zo = level6
    E, Z, res = m.solve_frequencies(wl.freqs, _params(m, path, 1e-9))
    assert all(solved(r, 1e-9) for r in res), res
    zr = Zo[1]
    dz = _rel(Z, zr)
    per_f = dz.max(axis=(1, 2))
    print(f"seed {seed} {path} tol 1e-9 vs converged oracle, max |dZ|/|Z| per frequency:",
          np.array2string(per_f, precision=1), flush=True)
    assert dz.max() <= 1e-5, per_f
    w = 2 * np.pi * wl.freqs[:, None, None]
    rho, rhor = np.abs(Z[..., 1:3]) ** 2 / (w * MU0), np.abs(zr[..., 1:3]) ** 2 / (w * MU0)
    assert (np.abs(rho - rhor) / rhor).max() <= 1e-3
    assert np.abs(np.degrees(np.angle(Z[..., 1:3] / zr[..., 1:3]))).max() <= 0.05
    scale = np.abs(zr[..., 1:3]).max(axis=(1, 2))[:, None, None]
    assert (np.abs(Z[..., [0, 3]] - zr[..., [0, 3]]) / scale).max() <= 1e-5


@pytest.mark.parametrize("path", PATHS)
def test_level6_reference_tolerance_within_oracle_spread(level6, path):
    seed, wl, m, Zo = level6
    E, Z, res = m.solve_frequencies(wl.freqs, _params(m, path, 1e-7))
    assert all(solved(r) for r in res), res
    spread = _rel(Zo[0], Zo[1]).max(axis=(1, 2))
    bar = np.maximum(1e-5, 3 * spread)
    dz = _rel(Z, Zo[0]).max(axis=(1, 2))
    own = _rel(Z, Zo[1]).max(axis=(1, 2))
    print(f"seed {seed} {path} tol 1e-7: vs oracle 1e-7", np.array2string(dz, precision=1),
          "| oracle spread", np.array2string(spread, precision=1),
          "| vs converged", np.array2string(own, precision=1), flush=True)
    assert (dz <= bar).all(), (dz, bar)
    assert (own <= bar).all(), (own, bar)
