"""Launch shapes and the autotuner (MI355X).

A launch shape only reassigns rows to GPU threads: every sweep split
(lanes per row) and grid cap must reproduce the oracle's SSOR in device row
order to 1e-12 relative (summation order only), and solves under tuned shapes
must meet the same BiCGStab parity as under the default shape.
"""
import numpy as np
import pytest

from maple3dmt_amd import MODE_FORWARD, MODE_SYSTEM, PREC_FP64, PREC_MIXED, configs
from maple3dmt_amd.native import NativeError, SolveParams
from conftest import solved
from oracle.oracle import Oracle, bicgstab_csr, ssor_csr
from test_gpu_parity import _permuted_csr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def block(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                          station_radius=5e3)
    m = configs.build_model(wl, ctx)
    o = Oracle(m, sigma_bg=wl.sigma_bg)
    o.set_sigma(m.array("sigma"))
    o.find_stations(wl.stations[:, :2])
    return wl, m, o


def _default_packed(w):
    return 1


def _default_wave(w):
    return 1 if w >= 8 else 0


def _default_tiled(w):
    return 1 if w >= 8 else 0


def _reset(dev):
    for prec in (PREC_FP64, PREC_MIXED):
        for w in (1, 2, 4, 8, 16, 32):
            dev.set_shape(prec, w, 1, 4096, 2048, 1, _default_packed(w), _default_wave(w), _default_tiled(w))


def test_shape_defaults_and_validation(block):
    _, m, _ = block
    dev = m.device
    for prec in (PREC_FP64, PREC_MIXED):
        for w in (1, 2, 4, 8, 16, 32):
            assert dev.get_shape(prec, w) == {"sweep_sub": 1, "sweep_cap": 4096, "row_cap": 2048,
                                              "sweep_chunk": 1, "packed": _default_packed(w),
                                              "wave": _default_wave(w), "tiled": _default_tiled(w)}
    for bad in ((3, 4096, 2048), (8, 4096, 2048), (1, 100, 2048), (1, 4096, 4096), (1, 0, 2048)):
        with pytest.raises(NativeError):
            dev.set_shape(PREC_FP64, 8, *bad)   # width 8: sub 8 gives 64 lanes per row
    with pytest.raises(NativeError):
        dev.get_shape(PREC_FP64, 3)
    with pytest.raises(NativeError):
        dev.set_shape(PREC_FP64, 8, 1, 4096, 2048, 3)   # chunk 3
    with pytest.raises(NativeError):
        dev.set_shape(PREC_FP64, 8, 1, 4096, 2048, 1, 2)   # packed 2
    with pytest.raises(NativeError):
        dev.set_shape(PREC_FP64, 4, 1, 4096, 2048, 1, 1, 1)   # wave below width 8
    with pytest.raises(NativeError):
        dev.set_shape(PREC_FP64, 4, 1, 4096, 2048, 1, 1, 0, 1)   # tiles below width 8
    dev.set_shape(PREC_FP64, 8, 4, 1024, 512, 4, 0, 0)
    assert dev.get_shape(PREC_FP64, 8) == {"sweep_sub": 4, "sweep_cap": 1024, "row_cap": 512,
                                           "sweep_chunk": 4, "packed": 0, "wave": 0, "tiled": 1}
    dev.set_shape(PREC_FP64, 8, 4, 1024, 512, 4, 0, 1)
    assert dev.get_shape(PREC_FP64, 8)["wave"] == 1
    dev.set_shape(PREC_FP64, 8, 4, 1024, 512, 4, 1)
    assert dev.get_shape(PREC_FP64, 8)["packed"] == 1
    _reset(dev)


@pytest.mark.parametrize("width", [1, 2, 4])
def test_every_sweep_split_matches_oracle_ssor(block, width):
    _, m, o = block
    dev = m.device
    w = 2 * np.pi * 0.5
    o.assemble(w)
    perm, _ = dev.permutation()
    ptr, col, val, real = _permuted_csr(o, perm)
    rng = np.random.default_rng(5)
    r = rng.standard_normal((width, m.n)) + 1j * rng.standard_normal((width, m.n))
    ref = np.zeros((width, m.n), np.complex128)
    for i in range(width):
        ref[i, real] = ssor_csr(ptr, col, val, r[i, real])
    try:
        for sub in (1, 2, 4, 8):
            if width * sub > 32:
                continue
            for cap, chunk, packed in ((8, 1, 1), (4096, 1, 0), (4096, 1, 1), (4096, 4, 1), (4096, 16, 0)):
                dev.set_shape(PREC_FP64, width, sub, cap, 2048, chunk, packed)
                print(f"width {width} sweep_sub {sub} sweep_cap {cap} sweep_chunk {chunk} packed {packed}",
                      flush=True)
                z = dev.ssor_apply(np.full(width, w), r)
                for i in range(width):
                    assert np.linalg.norm(z[i] - ref[i]) <= 1e-12 * np.linalg.norm(ref[i]), (sub, cap, i)
    finally:
        _reset(dev)


def test_autotuned_solves_keep_parity(block):
    wl, m, o = block
    dev = m.device
    w = 2 * np.pi * 1.0
    try:
        tuned = dev.autotune(np.full(32, w), widths=(1, 2), reps=8)
        for (prec, width), (ms, shape) in tuned.items():
            print(prec, width, ms, shape, flush=True)
            assert ms > 0
            assert width * shape["sweep_sub"] <= 32
        # FP64 BiCGStab against the oracle's natural-order iteration on the
        # device row order, as under the default shape.
        o.assemble(w)
        perm, _ = dev.permutation()
        ptr, col, val, real = _permuted_csr(o, perm)
        rng = np.random.default_rng(3)
        b = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
        tol = 1e-9
        x_dev, res = dev.solve(MODE_SYSTEM, [w], b[None, :], params=SolveParams(tol=tol, maxiter=3000))
        x_ref = np.zeros(m.n, np.complex128)
        xr, rr = bicgstab_csr(ptr, col, val, b[real], tol=tol, maxiter=3000)
        x_ref[real] = xr
        assert res[0]["converged"] and rr["converged"]
        assert abs(res[0]["iterations"] - rr["iterations"]) <= 2
        assert np.linalg.norm(x_dev[0] - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
        # Mixed-precision forward solve of both polarizations (complex64
        # kernels at width 2 under the tuned shape) against FP64.
        freqs = wl.freqs[:1]
        _, z64, r64 = m.solve_frequencies(freqs)
        _, z32, r32 = m.solve_frequencies(freqs, params=wl.fwd.solve_params(precision=PREC_MIXED))
        assert all(solved(r) for r in r64 + r32), r64 + r32
        dz = np.abs(z32[0][:, 1:3] - z64[0][:, 1:3]) / np.abs(z64[0][:, 1:3])
        assert dz.max() <= 1e-5
    finally:
        _reset(dev)
