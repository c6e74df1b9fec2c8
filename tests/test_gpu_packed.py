"""Packed SELL blocks (one 32-bit word per coupling: column and value-table
index) against the separate value / column blocks.

The packed layout changes only how a coupling's value is fetched (from a
table of the operator's distinct values), not the value nor the order of the
additions, so every result must be identical bit for bit: SpMV, SSOR apply,
and whole FP64 and mixed-precision solves (same iteration counts, same
fields).  Compared between two operators, one built without packed blocks
(M3D_PACKED_SELL=0), and within one operator by switching the launch shape's
layout choice.  The other GPU tests run the default choice (packed blocks at
every width) against the oracle.
"""
import os

import numpy as np
import pytest

from maple3dmt_amd import MODE_FORWARD, PREC_MIXED, configs

pytestmark = pytest.mark.gpu


def _build(ctx, wl, packed):
    old = os.environ.get("M3D_PACKED_SELL")
    os.environ["M3D_PACKED_SELL"] = "1" if packed else "0"
    try:
        return configs.build_model(wl, ctx)
    finally:
        if old is None:
            del os.environ["M3D_PACKED_SELL"]
        else:
            os.environ["M3D_PACKED_SELL"] = old


@pytest.fixture(scope="module")
def pair(ctx):
    wl = configs.random3d(5, min_level=3, max_level=6, n_freq=2, half_width=60e3, station_radius=8e3)
    mp = _build(ctx, wl, True)
    mu = _build(ctx, wl, False)
    yield wl, mp, mu
    mp.close()
    mu.close()


def test_format(pair):
    _, mp, mu = pair
    ip, iu = mp.device.info(), mu.device.info()
    print("packed", ip, "unpacked", iu, flush=True)
    assert 1 < ip["value_table"] <= 4096
    assert ip["column_bits"] >= int(np.ceil(np.log2(ip["n_padded"])))
    assert iu["value_table"] == 0 and iu["column_bits"] == 0
    assert {k: ip[k] for k in ("n_padded", "ncolours", "stored", "nnz")} == \
        {k: iu[k] for k in ("n_padded", "ncolours", "stored", "nnz")}


def _set_packed(dev, packed):
    for prec in (0, PREC_MIXED):
        for w in (1, 2, 4, 8, 16, 32):
            sh = dev.get_shape(prec, w)
            sh["packed"] = packed
            dev.set_shape(prec, w, **sh)


@pytest.fixture
def all_packed(pair):
    _, mp, _ = pair
    _set_packed(mp.device, 1)
    yield
    for prec in (0, PREC_MIXED):
        for w in (1, 2, 4, 8, 16, 32):
            sh = mp.device.get_shape(prec, w)
            sh["packed"] = 1
            mp.device.set_shape(prec, w, **sh)


@pytest.mark.parametrize("k", [1, 2, 5, 16])
def test_spmv_and_ssor_bitwise(pair, all_packed, k):
    _, mp, mu = pair
    rng = np.random.default_rng(k)
    om = 2 * np.pi * 10.0 ** rng.uniform(-3, 2, k)
    x = rng.standard_normal((k, mp.n)) + 1j * rng.standard_normal((k, mp.n))
    assert np.array_equal(mp.device.spmv(om, x), mu.device.spmv(om, x))
    assert np.array_equal(mp.device.ssor_apply(om, x), mu.device.ssor_apply(om, x))
    # The same operator with the layout switched by the launch shape.
    _set_packed(mp.device, 0)
    assert np.array_equal(mp.device.spmv(om, x), mu.device.spmv(om, x))
    assert np.array_equal(mp.device.ssor_apply(om, x), mu.device.ssor_apply(om, x))


@pytest.mark.parametrize("k", [8, 13, 16, 32])
def test_wavefront_shared_couplings_bitwise(pair, k):
    """From 8 systems the lanes of a row can load its couplings once, lane j
    entry j of a chunk, and pass them on through the LDS crossbar (launch
    shape wave = 1); the SpMV sums are the per-lane gather's bit for bit, from
    either block layout, and equal the twin operator built without packed
    blocks."""
    wl, mp, mu = pair
    rng = np.random.default_rng(100 + k)
    om = 2 * np.pi * 10.0 ** rng.uniform(-3, 2, k)
    x = rng.standard_normal((k, mp.n)) + 1j * rng.standard_normal((k, mp.n))
    w = 8 if k <= 8 else 16 if k <= 16 else 32
    dev = mp.device
    base = dev.get_shape(0, w)
    outs = {}
    try:
        for packed in (0, 1):
            for wave in (0, 1):
                dev.set_shape(0, w, base["sweep_sub"], base["sweep_cap"], base["row_cap"], base["sweep_chunk"],
                              packed, wave)
                outs[(packed, wave)] = dev.spmv(om, x)
    finally:
        dev.set_shape(0, w, **base)
    ref = outs[(0, 0)]
    for key, y in outs.items():
        assert np.array_equal(y, ref), key
    assert np.array_equal(mu.device.spmv(om, x), ref)


def _set_wave(dev, wave, packed=None):
    for prec in (0, PREC_MIXED):
        for w in (8, 16, 32):
            sh = dev.get_shape(prec, w)
            sh["wave"] = wave
            if packed is not None:
                sh["packed"] = packed
            dev.set_shape(prec, w, **sh)


@pytest.mark.parametrize("packed", [0, 1])
@pytest.mark.parametrize("k", [8, 13, 16])
def test_lane_shared_sweeps_and_solves_bitwise(pair, packed, k):
    """wave = 1 also shares couplings among a row's lanes in the SSOR sweeps
    and the BiCGStab products.  SSOR application and whole FP64 and mixed
    solves (systems that finish at different iterations, and 13 systems at
    width 16, whose three idle lanes pass couplings on without storing) equal
    the per-lane path's bit for bit."""
    wl, mp, _ = pair
    dev = mp.device
    saved = {(p, w): dev.get_shape(p, w) for p in (0, PREC_MIXED) for w in (8, 16, 32)}
    rng = np.random.default_rng(200 + k)
    om = 2 * np.pi * 10.0 ** rng.uniform(-3, 2, k)
    x = rng.standard_normal((k, mp.n)) + 1j * rng.standard_normal((k, mp.n))
    rhs = np.stack([mp.rhs_total(mp.bc_values(w, q % 2)) for q, w in enumerate(om)])
    aux = np.stack([mp.primary_field(w, q % 2) for q, w in enumerate(om)])
    outs = []
    try:
        for wave in (0, 1):
            _set_wave(dev, wave, packed)
            o = {"ssor": dev.ssor_apply(om, x)}
            for precision in (0, PREC_MIXED):
                prm = mp.fwd_params.solve_params(precision=precision, inner_tol=1e-3 if precision else 0.0)
                xs, res = dev.solve(MODE_FORWARD, om, rhs, aux, params=prm)
                o[precision] = (xs, [r["iterations"] for r in res], [r["info"] for r in res])
            outs.append(o)
    finally:
        for (p, w), sh in saved.items():
            dev.set_shape(p, w, **sh)
    print("iterations", outs[0][0][1], outs[1][0][1], outs[0][PREC_MIXED][1], flush=True)
    assert np.array_equal(outs[0]["ssor"], outs[1]["ssor"])
    for precision in (0, PREC_MIXED):
        assert outs[0][precision][1] == outs[1][precision][1]
        assert outs[0][precision][2] == outs[1][precision][2]
        assert np.array_equal(outs[0][precision][0], outs[1][precision][0])


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_solve_bitwise(pair, all_packed, precision):
    wl, mp, mu = pair
    prm = mp.fwd_params.solve_params(precision=precision, inner_tol=1e-3 if precision else 0.0)
    omegas = np.repeat(2 * np.pi * wl.freqs, 2)
    out = []
    for m in (mp, mu):
        rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
        aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(omegas)])
        x, res = m.device.solve(MODE_FORWARD, omegas, rhs, aux, params=prm)
        out.append((x, [r["iterations"] for r in res], [r["info"] for r in res]))
    print("iterations", out[0][1], out[1][1], flush=True)
    assert out[0][1] == out[1][1]
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.fixture(scope="module")
def unfused(ctx, pair):
    """The packed operator's twin with the last colour's backward step in a
    launch of its own (M3D_FUSE_LAST_COLOUR=0)."""
    wl = pair[0]
    old = os.environ.get("M3D_FUSE_LAST_COLOUR")
    os.environ["M3D_FUSE_LAST_COLOUR"] = "0"
    try:
        m = configs.build_model(wl, ctx)
    finally:
        if old is None:
            del os.environ["M3D_FUSE_LAST_COLOUR"]
        else:
            os.environ["M3D_FUSE_LAST_COLOUR"] = old
    yield m
    m.close()


@pytest.mark.parametrize("k", [1, 2, 16])
def test_fused_last_colour_bitwise(pair, unfused, k):
    """Writing z = y of the last colour from its forward launch gives the
    separate backward launch's values bit for bit, in the SSOR apply and in
    whole solves."""
    wl, mp, _ = pair
    rng = np.random.default_rng(10 + k)
    om = 2 * np.pi * 10.0 ** rng.uniform(-3, 2, k)
    x = rng.standard_normal((k, mp.n)) + 1j * rng.standard_normal((k, mp.n))
    assert np.array_equal(mp.device.ssor_apply(om, x), unfused.device.ssor_apply(om, x))
    if k == 2:
        omegas = np.repeat(2 * np.pi * wl.freqs[:1], 2)
        out = []
        for m in (mp, unfused):
            rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
            aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(omegas)])
            xs, res = m.device.solve(MODE_FORWARD, omegas, rhs, aux, params=m.fwd_params.solve_params())
            out.append((xs, [r["iterations"] for r in res]))
        assert out[0][1] == out[1][1]
        assert np.array_equal(out[0][0], out[1][0])


def _raw_operator(ctx, n, values, seed):
    """A device operator from a random sparse CSR (a symmetric pattern of
    about 12 couplings per row, no constrained rows) whose couplings take
    values(rng, nnz)."""
    import ctypes as C
    import scipy.sparse as sp
    from maple3dmt_amd.native import DeviceOperator, OperatorDescC, _check_dev, lib
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), 6)
    cols = rng.integers(0, n, rows.size)
    keep = rows != cols
    A = sp.csr_matrix((np.ones(keep.sum()), (rows[keep], cols[keep])), shape=(n, n))
    A = (A + A.T).tocsr()                 # a symmetric pattern, as the colouring expects
    A.sum_duplicates()
    A.sort_indices()
    A.data = values(rng, A.nnz)
    ptr = A.indptr.astype(np.int32)
    col = A.indices.astype(np.int32)
    val = np.ascontiguousarray(A.data, np.float64)
    diag = rng.uniform(4.0, 8.0, n)
    me = rng.uniform(0.1, 1.0, n)
    zero = np.zeros(n, np.uint8)
    keep_alive = (ptr, col, val, diag, me, zero)
    d = OperatorDescC(n, ptr.ctypes.data_as(C.POINTER(C.c_int)), col.ctypes.data_as(C.POINTER(C.c_int)),
                      val.ctypes.data_as(C.POINTER(C.c_double)), diag.ctypes.data_as(C.POINTER(C.c_double)),
                      zero.ctypes.data_as(C.POINTER(C.c_ubyte)), zero.ctypes.data_as(C.POINTER(C.c_ubyte)),
                      0, None)
    h = C.c_void_p()
    _check_dev(lib().m3d_operator_create(ctx.handle, C.byref(d), C.byref(h)))
    _check_dev(lib().m3d_operator_set_mass(h, me.ctypes.data_as(C.c_void_p)))
    return DeviceOperator(h, n=n), A, diag, me, keep_alive


@pytest.mark.parametrize("distinct", [50, None])
def test_packing_on_general_matrices(ctx, distinct):
    """Couplings drawn from 50 values are packed (a table of 51 with the
    padding zero); random couplings (over 4096 distinct values) are not.
    Either way the SpMV at 1 and 4 systems matches SciPy's product."""
    from maple3dmt_amd.native import _check_dev, lib
    if distinct:
        def values(rng, k):
            return rng.choice(rng.standard_normal(distinct), k)
    else:
        def values(rng, k):
            return rng.standard_normal(k)
    dev, A, diag, me, _keep = _raw_operator(ctx, 20000, values, 5)
    try:
        info = dev.info()
        if distinct:
            assert info["value_table"] == distinct + 1
        else:
            assert info["value_table"] == 0
        rng = np.random.default_rng(6)
        for k in (1, 4):
            om = 2 * np.pi * rng.uniform(0.1, 10.0, k)
            x = rng.standard_normal((k, A.shape[0])) + 1j * rng.standard_normal((k, A.shape[0]))
            y = dev.spmv(om, x)
            for i in range(k):
                ref = A @ x[i] + (diag + 1j * om[i] * me) * x[i]
                assert np.linalg.norm(y[i] - ref) <= 1e-13 * np.linalg.norm(ref)
    finally:
        _check_dev(lib().m3d_operator_destroy(dev.handle))


@pytest.fixture(scope="module")
def device_order(ctx, pair):
    """The packed operator's twin whose plain SpMV visits row blocks in device
    (colour) order (M3D_SPMV_ORDER=0) instead of the spatial block order."""
    wl = pair[0]
    old = os.environ.get("M3D_SPMV_ORDER")
    os.environ["M3D_SPMV_ORDER"] = "0"
    try:
        m = configs.build_model(wl, ctx)
    finally:
        if old is None:
            del os.environ["M3D_SPMV_ORDER"]
        else:
            os.environ["M3D_SPMV_ORDER"] = old
    yield m
    m.close()


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16, 32])
def test_spmv_block_order_bitwise(pair, device_order, k):
    """Visiting the row blocks in spatial order (each XCD walking one
    contiguous part) changes which rows run together, not a row's sum: the
    products are identical bit for bit at every batch width (a row left out
    would keep a stale workspace value and differ from the twin's, which
    writes every row in device order)."""
    _, mp, _ = pair
    rng = np.random.default_rng(40 + k)
    om = 2 * np.pi * 10.0 ** rng.uniform(-3, 2, k)
    x = rng.standard_normal((k, mp.n)) + 1j * rng.standard_normal((k, mp.n))
    y = mp.device.spmv(om, x)
    assert np.all(np.isfinite(y))
    assert np.array_equal(y, device_order.device.spmv(om, x))
