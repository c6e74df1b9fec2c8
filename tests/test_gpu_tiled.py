"""The plain product over LDS-staged row tiles (launch shape tiled = 1, widths
8-32) against the product that gathers x from device memory (tiled = 0).

A tile stages the x rows its couplings read into LDS and every coupling reads
x from there; a row adds the same entries in the same order either way, and
its padding adds 0 * 0, so the products are equal bit for bit: on random
vectors at every tiled width (idle lanes included), and through whole forward
solves, whose right-hand side b - A E_pri the product forms.  The tiled
product also meets the oracle's matvec to 1e-12 (summation order only) on the
heterogeneous levels 3-5 model.
"""
import numpy as np
import pytest

from maple3dmt_amd import MODE_FORWARD, PREC_MIXED, configs
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(ctx):
    wl = configs.random3d(5, min_level=3, max_level=6, n_freq=2, half_width=60e3, station_radius=8e3)
    m = configs.build_model(wl, ctx)
    yield wl, m
    m.close()


def _width(k):
    return 8 if k <= 8 else 16 if k <= 16 else 32


def _with_tiled(dev, tiled, fn, widths=(8, 16, 32)):
    saved = {(p, w): dev.get_shape(p, w) for p in (0, PREC_MIXED) for w in widths}
    try:
        for (p, w), sh in saved.items():
            dev.set_shape(p, w, **{**sh, "tiled": tiled})
        return fn()
    finally:
        for (p, w), sh in saved.items():
            dev.set_shape(p, w, **sh)


def test_default_shape_is_tiled(model):
    _, m = model
    for p in (0, PREC_MIXED):
        for w in (1, 2, 4):
            assert m.device.get_shape(p, w)["tiled"] == 0
        for w in (8, 16, 32):
            assert m.device.get_shape(p, w)["tiled"] == 1


@pytest.mark.parametrize("k", [5, 8, 13, 16, 23, 32])
def test_tiled_spmv_bitwise(model, k):
    _, m = model
    rng = np.random.default_rng(300 + k)
    om = 2 * np.pi * 10.0 ** rng.uniform(-3, 2, k)
    x = rng.standard_normal((k, m.n)) + 1j * rng.standard_normal((k, m.n))
    dev = m.device
    y_tiled = _with_tiled(dev, 1, lambda: dev.spmv(om, x))
    y_plain = _with_tiled(dev, 0, lambda: dev.spmv(om, x))
    assert np.all(np.isfinite(y_tiled))
    assert np.array_equal(y_tiled, y_plain)


def test_tiled_forward_solves_bitwise(model):
    wl, m = model
    dev = m.device
    om = np.repeat(2 * np.pi * np.logspace(-3, 2, 8), 2)
    rhs = np.stack([m.rhs_total(m.bc_values(w, q % 2)) for q, w in enumerate(om)])
    aux = np.stack([m.primary_field(w, q % 2) for q, w in enumerate(om)])
    out = []
    for tiled in (1, 0):
        for precision in (0, PREC_MIXED):
            prm = m.fwd_params.solve_params(precision=precision, inner_tol=1e-3 if precision else 0.0)
            xs, res = _with_tiled(dev, tiled, lambda: dev.solve(MODE_FORWARD, om, rhs, aux, params=prm))
            out.append((xs, [r["iterations"] for r in res], [r["info"] for r in res]))
    print("iterations", out[0][1], out[1][1], flush=True)
    for a, b in ((out[0], out[2]), (out[1], out[3])):
        assert a[1] == b[1]
        assert a[2] == b[2]
        assert np.array_equal(a[0], b[0])


def test_tiled_spmv_matches_oracle(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    try:
        o = Oracle(m, sigma_bg=wl.sigma_bg)
        o.set_sigma(m.array("sigma"))
        rng = np.random.default_rng(11)
        for w in (2 * np.pi * 1e-3, 2 * np.pi * 10.0):
            o.assemble(w)
            xs = rng.standard_normal((16, m.n)) + 1j * rng.standard_normal((16, m.n))
            assert m.device.get_shape(0, 16)["tiled"] == 1
            ys = m.device.spmv(np.full(16, w), xs)
            for x, y in zip(xs, ys):
                ref = o.spmv(x)
                assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    finally:
        m.close()


@pytest.mark.parametrize("distinct", [50, 4000])
def test_tiled_spmv_general_matrix(ctx, distinct):
    """A random sparse pattern (no spatial order, ~13 staged rows per row) whose
    couplings take `distinct` values: with 4000 the value table fills 32 KB of
    the tile's LDS and the tiles shrink to the rest.  The tiled product at 8,
    16 and 32 systems matches SciPy's and the gathering product bit for bit."""
    from maple3dmt_amd.native import _check_dev, lib
    from test_gpu_packed import _raw_operator

    def values(rng, k):
        return rng.choice(rng.standard_normal(distinct), k)
    dev, A, diag, me, _keep = _raw_operator(ctx, 12000, values, 9)
    try:
        assert dev.info()["value_table"] == distinct + 1
        rng = np.random.default_rng(10)
        for k in (8, 16, 32):
            om = 2 * np.pi * rng.uniform(0.1, 10.0, k)
            x = rng.standard_normal((k, A.shape[0])) + 1j * rng.standard_normal((k, A.shape[0]))
            assert dev.get_shape(0, k)["tiled"] == 1
            y = dev.spmv(om, x)
            y0 = _with_tiled(dev, 0, lambda: dev.spmv(om, x))
            assert np.array_equal(y, y0)
            for i in range(k):
                ref = A @ x[i] + (diag + 1j * om[i] * me) * x[i]
                assert np.linalg.norm(y[i] - ref) <= 1e-13 * np.linalg.norm(ref)
    finally:
        _check_dev(lib().m3d_operator_destroy(dev.handle))
