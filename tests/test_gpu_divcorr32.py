"""The FP32 DivCorr PCG of the mixed-precision path against the FP64 one.

With M3D_DIVCORR_FP32=1, precision = mixed and a DivCorr PCG tolerance of at
least 1e-5 (the fast path's 1e-4), the DivCorr steps between the interleaved
bursts run their PCG in FP32 (complex64 node vectors, FP32 Laplacian, double
reductions); by default they stay FP64.  The field E, the divergence that
stops DivCorr and the true residual that stops the solve stay FP64, so the
result meets the same tolerance: the impedances of both forms agree within
the 1e-5 parity bar (the solve tolerance decides their difference, not the
PCG's precision).  Everything else - the FP64 paths, PCG tolerances below
1e-5 - ignores the switch bit for bit.
"""
import numpy as np
import pytest

from maple3dmt_amd import configs
from maple3dmt_amd.native import PREC_MIXED
from conftest import solved

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3,
                          maxiter=40000)
    m = configs.build_model(wl, ctx)
    yield wl, m
    m.close()


def _solve(m, wl, prm, fp32, monkeypatch):
    monkeypatch.setenv("M3D_DIVCORR_FP32", "1" if fp32 else "0")
    return m.solve_frequencies(wl.freqs, prm)


def _rel(z, zr):
    return np.abs(z[..., 1:3] - zr[..., 1:3]) / np.abs(zr[..., 1:3])


@pytest.mark.parametrize("tol", [1e-7, 1e-9])
def test_fast_path_fp32_divcorr_matches_fp64(model, tol, monkeypatch):
    wl, m = model
    prm = m.fwd_params.fast_params(tol=tol)
    prm.maxiter = 40000
    assert prm.precision == PREC_MIXED and prm.divcorr_pcg_tol == 1e-4
    E32, Z32, r32 = _solve(m, wl, prm, True, monkeypatch)
    E64, Z64, r64 = _solve(m, wl, prm, False, monkeypatch)
    assert all(solved(r, tol) for r in r32), r32
    assert all(solved(r, tol) for r in r64), r64
    # The FP32 PCG is the one that ran: the iterates differ.
    assert not np.array_equal(E32, E64)
    err = _rel(Z32, Z64)
    print("iterations fp32", [r["iterations"] for r in r32], "fp64", [r["iterations"] for r in r64],
          "max |dZ|/|Z|", float(err.max()), flush=True)
    assert err.max() <= 1e-5


@pytest.mark.parametrize("path", ["fp64_interleaved", "mixed_pcg_1e-6", "mixed_post_hoc"])
def test_other_paths_ignore_switch(model, path, monkeypatch):
    wl, m = model
    if path == "fp64_interleaved":
        prm = m.fwd_params.fast_params(precision=0)
    elif path == "mixed_pcg_1e-6":
        prm = m.fwd_params.fast_params()
        prm.divcorr_pcg_tol = 1e-6
    else:  # mixed precision, the reference's post-hoc DivCorr at 1e-10
        prm = m.fwd_params.solve_params(precision=PREC_MIXED, inner_tol=1e-3)
    prm.maxiter = 40000
    Ea, Za, ra = _solve(m, wl, prm, True, monkeypatch)
    Eb, Zb, rb = _solve(m, wl, prm, False, monkeypatch)
    assert all(solved(r) for r in ra), ra
    assert [r["iterations"] for r in ra] == [r["iterations"] for r in rb]
    assert np.array_equal(Ea, Eb)
    assert np.array_equal(Za, Zb)
