"""Writes the oracle's impedances on the level-6 random models used by the
GPU parity tests (tests/test_gpu_parity_level6.py).

    python tests/golden/make_level6_fixtures.py [seed ...]

For each seed: the random 3-D model of configs.random3d(seed, max_level=6,
half_width=100 km, station_radius=20 km) with its 8 frequencies
(1e-3..1e2 Hz) and 121 stations, solved by the CPU oracle (natural-order
SSOR-BiCGStab, post-hoc DivCorr: forward_solver_fv.cpp:301-385) twice:
at the reference's tolerance 1e-7 and converged at 1e-10.  Stored: both
impedance sets, the iteration counts, and a fingerprint of the mesh (edge,
cell and node counts and sums of the edge coordinates) that the GPU test
checks before comparing, so the fixture can only be used on the same mesh.
Runs on the CPU in a few minutes per seed.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from maple3dmt_amd import configs  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

TOLS = (1e-7, 1e-10)


def workload(seed):
    return configs.random3d(seed, max_level=6, half_width=100e3, station_radius=20e3, maxiter=200000)


def fingerprint(m):
    return np.array([m.n, m.nc, m.nn, m.array("edge_x").sum(), m.array("edge_y").sum(),
                     m.array("edge_z").sum()])


def make(seed):
    wl = workload(seed)
    m = configs.build_model(wl)
    Z = np.zeros((len(TOLS), len(wl.freqs), len(wl.stations), 4), np.complex128)
    its = np.zeros((len(TOLS), len(wl.freqs), 2), np.int64)
    for ti, tol in enumerate(TOLS):
        fwd = wl.fwd
        fwd.bicgstab_tol = tol
        o = Oracle(m, fwd=fwd, sigma_bg=wl.sigma_bg)
        o.set_sigma(m.array("sigma"))
        o.find_stations(wl.stations[:, :2])
        for fi, f in enumerate(wl.freqs):
            t = time.time()
            _, _, z, res = o.solve_frequency(f)
            assert res[0]["converged"] and res[1]["converged"], res
            Z[ti, fi] = z
            its[ti, fi] = [res[0]["iterations"], res[1]["iterations"]]
            print(f"seed {seed} tol {tol:.0e} f {f:.3g} Hz: {its[ti, fi]} its, {time.time() - t:.1f}s",
                  flush=True)
    out = os.path.join(HERE, f"level6_seed{seed}.npz")
    np.savez_compressed(out, Z=Z, tols=np.array(TOLS), freqs=wl.freqs, iterations=its,
                        fingerprint=fingerprint(m))
    print("wrote", out, flush=True)


if __name__ == "__main__":
    for s in (sys.argv[1:] or ["50001", "50003"]):
        make(int(s))
