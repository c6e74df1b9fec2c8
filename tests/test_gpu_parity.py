"""Device solve vs the CPU oracle on the same inputs (MI355X).

Tolerances (floating point, FP64 unless stated):
  * SpMV, SSOR apply, DivCorr: relative 1e-12 (summation order only).
  * BiCGStab on the colour-ordered matrix: the oracle's natural-order SSOR on
    the device's row order is the device's multicolour SSOR, so iteration
    counts must agree within 1 and solutions to 1e-8 relative.
  * Forward impedance: |dZ|/|Z| <= 1e-5, rho_a within 0.1 %, phase within
    0.05 degrees against the oracle's solve.
"""
import numpy as np
import pytest

from maple3dmt_amd import (MODE_FORWARD, MODE_INNER, MODE_INTERLEAVED, MODE_SYSTEM, PREC_MIXED,
                           configs)
from maple3dmt_amd.native import SolveParams
from conftest import solved
from oracle.oracle import Oracle, bicgstab_csr, ssor_csr

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _rng(seed):
    return np.random.default_rng(seed)


def _cvec(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.fixture(scope="module")
def halfspace(ctx):
    wl = configs.halfspace()
    m = configs.build_model(wl, ctx)
    o = Oracle(m)
    o.set_sigma(m.array("sigma"))
    o.find_stations(wl.stations[:, :2])
    return wl, m, o


@pytest.fixture(scope="module")
def block(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                          station_radius=5e3)
    m = configs.build_model(wl, ctx)
    o = Oracle(m, sigma_bg=wl.sigma_bg)
    o.set_sigma(m.array("sigma"))
    o.find_stations(wl.stations[:, :2])
    return wl, m, o


def _permuted_csr(o, perm):
    """The oracle's eliminated A in the device row order (padding dropped)."""
    ptr, col, val = o.csr("A")
    real = perm[perm >= 0]
    n = len(real)
    inv = np.full(len(ptr) - 1, -1)
    inv[real] = np.arange(n)
    rows, cols, vals = [], [], []
    for new, old in enumerate(real):
        for k in range(ptr[old], ptr[old + 1]):
            rows.append(new)
            cols.append(inv[col[k]])
            vals.append(val[k])
    rows, cols, vals = np.array(rows), np.array(cols), np.array(vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    p = np.zeros(n + 1, np.int32)
    np.add.at(p, rows + 1, 1)
    return np.cumsum(p).astype(np.int32), cols.astype(np.int32), vals, real


@pytest.mark.parametrize("which", ["halfspace", "block"])
def test_spmv_matches_oracle(which, request):
    wl, m, o = request.getfixturevalue(which)
    rng = _rng(1)
    for w in (2 * np.pi * 1e-3, 2 * np.pi * 1.0):
        o.assemble(w)
        xs = np.stack([_cvec(rng, m.n) for _ in range(3)])
        ys = m.device.spmv(np.full(3, w), xs)
        for x, y in zip(xs, ys):
            ref = o.spmv(x)
            assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("which", ["halfspace", "block"])
def test_ssor_matches_natural_order_in_device_order(which, request):
    wl, m, o = request.getfixturevalue(which)
    w = 2 * np.pi * 0.5
    o.assemble(w)
    perm, cs = m.device.permutation()
    ptr, col, val, real = _permuted_csr(o, perm)
    rng = _rng(2)
    r = _cvec(rng, m.n)
    z_dev = m.device.ssor_apply([w], r)[0]
    z_ref = np.zeros(m.n, np.complex128)
    z_ref[real] = ssor_csr(ptr, col, val, r[real])
    assert np.linalg.norm(z_dev - z_ref) <= 1e-12 * np.linalg.norm(z_ref)


@pytest.mark.parametrize("which", ["halfspace", "block"])
def test_bicgstab_iterates_match_oracle(which, request):
    wl, m, o = request.getfixturevalue(which)
    w = 2 * np.pi * 1.0
    o.assemble(w)
    perm, cs = m.device.permutation()
    ptr, col, val, real = _permuted_csr(o, perm)
    rng = _rng(3)
    b = _cvec(rng, m.n)
    tol = 1e-9
    x_dev, res = m.device.solve(MODE_SYSTEM, [w], b[None, :], params=SolveParams(tol=tol, maxiter=3000))
    x_ref = np.zeros(m.n, np.complex128)
    xr, rr = bicgstab_csr(ptr, col, val, b[real], tol=tol, maxiter=3000)
    x_ref[real] = xr
    assert res[0]["converged"] and rr["converged"]
    assert abs(res[0]["iterations"] - rr["iterations"]) <= 1
    assert np.linalg.norm(x_dev[0] - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    assert res[0]["true_residual"] < 10 * tol


def _z_errors(z, zr, w):
    dz = np.abs(z - zr) / np.abs(zr)
    rho = np.abs(z) ** 2 / (w * MU0)
    rhor = np.abs(zr) ** 2 / (w * MU0)
    drho = np.abs(rho - rhor) / rhor
    dph = np.abs(np.degrees(np.angle(z / zr)))
    return dz, drho, dph


@pytest.mark.parametrize("which", ["halfspace", "block"])
def test_forward_impedance_parity(which, request):
    wl, m, o = request.getfixturevalue(which)
    freqs = wl.freqs
    E, Z, res = m.solve_frequencies(freqs)
    assert all(solved(r) for r in res), res
    for fi, f in enumerate(freqs):
        w = 2 * np.pi * f
        e1, e2, zr, _ = o.solve_frequency(f)
        # Off-diagonal impedances carry the response; diagonal ones are ~0 in
        # 1-D structures and are compared in absolute terms.
        dz, drho, dph = _z_errors(Z[fi][:, 1:3], zr[:, 1:3], w)
        assert dz.max() <= 1e-5, (f, dz.max())
        assert drho.max() <= 1e-3
        assert dph.max() <= 0.05
        scale = np.abs(zr[:, 1:3]).max()
        assert np.abs(Z[fi][:, [0, 3]] - zr[:, [0, 3]]).max() <= 1e-5 * scale


def test_forward_mixed_precision_parity(block):
    wl, m, o = block
    prm = m.fwd_params.solve_params(precision=PREC_MIXED, inner_tol=1e-3)
    E, Z, res = m.solve_frequencies(wl.freqs, prm)
    assert all(solved(r) for r in res), res
    for fi, f in enumerate(wl.freqs):
        _, _, zr, _ = o.solve_frequency(f)
        dz, drho, dph = _z_errors(Z[fi][:, 1:3], zr[:, 1:3], 2 * np.pi * f)
        assert dz.max() <= 1e-5 and drho.max() <= 1e-3 and dph.max() <= 0.05


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_forward_interleaved_divcorr_parity(block, precision):
    # E_sec by interleaved BiCGStab + DivCorr bursts instead of one full
    # BiCGStab: the same system to the same tolerance, so Z meets the same bar.
    wl, m, o = block
    prm = m.fwd_params.solve_params(precision=precision, inner_tol=1e-3 if precision else 0.0)
    prm.interleaved_forward = 1
    E, Z, res = m.solve_frequencies(wl.freqs, prm)
    assert all(solved(r) for r in res), res
    for fi, f in enumerate(wl.freqs):
        _, _, zr, _ = o.solve_frequency(f)
        dz, drho, dph = _z_errors(Z[fi][:, 1:3], zr[:, 1:3], 2 * np.pi * f)
        assert dz.max() <= 1e-5 and drho.max() <= 1e-3 and dph.max() <= 0.05, (f, dz.max())


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_fast_path_parity(block, precision):
    # The fast path's bursts of 100 iterations with a DivCorr PCG stopped at
    # 1e-4 instead of 1e-10: the same FP64 true-residual tolerance, so Z meets
    # the same bar against the oracle.
    wl, m, o = block
    prm = m.fwd_params.fast_params(precision=precision)
    assert prm.divcorr_pcg_tol == 1e-4
    E, Z, res = m.solve_frequencies(wl.freqs, prm)
    assert all(solved(r) for r in res), res
    for fi, f in enumerate(wl.freqs):
        _, _, zr, _ = o.solve_frequency(f)
        dz, drho, dph = _z_errors(Z[fi][:, 1:3], zr[:, 1:3], 2 * np.pi * f)
        assert dz.max() <= 1e-5 and drho.max() <= 1e-3 and dph.max() <= 0.05, (f, dz.max())


def test_divcorr_matches_oracle(block):
    wl, m, o = block
    w = 2 * np.pi * wl.freqs[0]
    o.assemble(w)
    rng = _rng(5)
    E = _cvec(rng, m.n)
    phi0 = 1e-3 * _cvec(rng, m.nn)
    for p in (None, phi0):
        ref = o.divcorr(E, p)
        got = m.device.divcorr([w], E[None, :], None if p is None else p[None, :],
                               m.fwd_params.solve_params())[0]
        assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)


def test_adjoint_interleaved_matches_oracle(block):
    wl, m, o = block
    f = wl.freqs[-1]
    w = 2 * np.pi * f
    E, Z, _ = m.solve_frequencies([f])
    e1, e2, zr, _ = o.solve_frequency(f)
    m.extract_impedance(w, E[0], E[1])
    rng = _rng(6)
    wr = np.conj(_cvec(rng, zr.size).reshape(zr.shape) * np.abs(zr))
    r1, r2 = m.adjoint_rhs(w, wr)
    q1, q2 = o.adjoint_rhs(wr)
    assert np.linalg.norm(r1 - q1) <= 1e-6 * np.linalg.norm(q1)
    lam, res = m.adjoint_solve([w, w], np.stack([r1, r2]))
    for k, rhs in enumerate((q1, q2)):
        lr, rr = o.adjoint(rhs)
        assert res[k]["converged"] == rr["converged"]
        assert res[k]["true_residual"] < 1e-6
        assert np.linalg.norm(lam[k] - lr) <= 1e-4 * np.linalg.norm(lr)
    g = m.sensitivity(w, E[0], lam[0]) + m.sensitivity(w, E[1], lam[1])
    gr = o.sensitivity(e1, o.adjoint(q1)[0]) + o.sensitivity(e2, o.adjoint(q2)[0])
    assert np.linalg.norm(g - gr) <= 1e-3 * np.linalg.norm(gr)


def test_batched_matches_single(block):
    # Batch width changes the block decomposition of the reductions, so the
    # iterates differ by rounding; both stop at the 1e-7 residual.
    wl, m, o = block
    E2, Z2, _ = m.solve_frequencies(wl.freqs)
    for fi, f in enumerate(wl.freqs):
        E1, Z1, _ = m.solve_frequencies([f])
        rel = float(np.abs(Z1[0] - Z2[fi]).max() / np.abs(Z2[fi]).max())
        assert rel <= 1e-6, rel
