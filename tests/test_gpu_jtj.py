"""Gauss-Newton product J^T W^2 J x (ImplicitJtJOperatorFV::Mult,
implicit_jtj_fv.cpp:78-192) on the device.

* Perturbation field.  In a uniform medium on a uniform mesh the reference's
  harmonic-mean L^3 mass derivative (build_perturbation_rhs,
  forward_solver_fv.cpp:945-984) equals the derivative of the edge mass
  (A_dual = L^2, harmonic = arithmetic mean), so the perturbation solve must
  reproduce the central difference of the forward field along x.  Both sides
  run without divergence correction, so they are plain linear solves
  E(sigma) = A(sigma)^-1 b and dE = A^-1 P(sigma x); tolerance 1e-5 relative
  (the difference step is 1e-4, the solves converge to 1e-11).
* The batched operator against the same product evaluated in the reference's
  order (frequency by frequency, polarisation by polarisation, one system per
  solve, host buffers), with inner solves tight enough (1e-9) that the two
  agree to 1e-5 in norm.
"""
import numpy as np
import pytest

from maple3dmt_amd import configs
from maple3dmt_amd.jtj import GaussNewton, matvec_host
from maple3dmt_amd.native import ForwardParams, MeshParams, Model

pytestmark = pytest.mark.gpu


def _uniform_model(ctx):
    mesh = MeshParams(domain=(-20e3, 20e3, -20e3, 20e3, -20e3, 20e3), min_level=4, max_level=4,
                      sigma_bg=0.01)
    st = configs.station_grid(3, 2500.0)
    fwd = ForwardParams(bicgstab_tol=1e-11, bicgstab_maxiter=20000, div_correction=False)
    m = Model(mesh, st, np.array([1.0]), fwd, ctx)
    m.set_sigma(np.full(m.nc, 0.01))
    m.find_stations(st[:, :2])
    return m


def test_perturbation_field_matches_central_difference(ctx):
    m = _uniform_model(ctx)
    freqs = np.array([1.0])
    # x on earth cells whose edges are all free (the reference's perturbation
    # right-hand side is not zeroed on Dirichlet rows, where a difference of
    # forward fields is zero by construction).
    ptr, ce = m.array("cell_edges_ptr"), m.array("cell_edges")
    cons = m.array("is_bc").astype(bool)
    cz = m.array("cell_cz")
    free = np.array([not cons[ce[ptr[c]:ptr[c + 1]]].any() for c in range(m.nc)])
    rng = np.random.default_rng(6)
    x = np.where(free & (cz < 0), rng.uniform(0.5, 1.5, m.nc), 0.0)
    assert np.count_nonzero(x) > 100
    gn = GaussNewton(m, freqs, np.ones((1, m.count("stations"), 4)), inner_tol=1e-11)
    _, d_de = gn.jvp(x)
    de = d_de.cpu().numpy()
    s0 = m.array("sigma").copy()
    eps = 1e-4
    try:
        m.set_sigma(s0 * np.exp(eps * x))
        ep, _, rp = m.solve_frequencies(freqs)
        m.set_sigma(s0 * np.exp(-eps * x))
        em, _, rm = m.solve_frequencies(freqs)
    finally:
        m.set_sigma(s0)
    assert all(r["converged"] for r in rp + rm)
    fd = (ep - em) / (2 * eps)
    for p in range(2):
        err = np.linalg.norm(de[p] - fd[p]) / np.linalg.norm(fd[p])
        print("polarisation", p + 1, "relative difference", err, flush=True)
        assert err < 1e-5


def test_device_perturbation_rhs_matches_host(ctx):
    """m3d_perturbation_rhs_device equals the host build_perturbation_rhs
    restatement bit for bit (heterogeneous model, random fields and dsigma
    with zeros)."""
    import torch
    dev = torch.device("cuda:0")
    wl = configs.random3d(8, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    rng = np.random.default_rng(9)
    e = rng.standard_normal((3, m.n)) + 1j * rng.standard_normal((3, m.n))
    ds = rng.standard_normal(m.nc) * m.array("sigma")
    ds[rng.uniform(size=m.nc) < 0.3] = 0.0
    om = np.array([2 * np.pi * 0.01, 2 * np.pi * 1.0, 2 * np.pi * 30.0])
    ref = np.stack([m.perturbation_rhs(w, e[k], ds) for k, w in enumerate(om)])
    d_e = torch.from_numpy(e).to(dev)
    d_ds = torch.from_numpy(ds).to(dev)
    d_r = torch.empty_like(d_e)
    torch.cuda.synchronize()
    m.device.perturbation_rhs_device(om, d_e.data_ptr(), d_ds.data_ptr(), d_r.data_ptr())
    assert np.array_equal(d_r.cpu().numpy(), ref)


def test_batched_product_matches_reference_order(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    ns = m.count("stations")
    err = np.full((len(wl.freqs), ns, 4), 0.05)
    err[:, :, [0, 3]] = 0.0               # diagonal components carry no weight
    earth = m.array("sigma") > 1e-6
    x = np.where(earth, np.random.default_rng(8).standard_normal(m.nc), 0.0)
    gn = GaussNewton(m, wl.freqs, err, inner_tol=1e-9)
    y = gn.matvec(x)
    y_ref = matvec_host(m, wl.freqs, err, x, inner_tol=1e-9)
    assert np.linalg.norm(y - y_ref) <= 1e-5 * np.linalg.norm(y_ref)
    assert np.count_nonzero(y[~earth]) <= np.count_nonzero(y)   # air cells carry little


@pytest.mark.parametrize("h_cache", ["reference", "per_frequency"])
def test_jvp_and_product_match_oracle(ctx, h_cache):
    """delta Z of every (frequency, polarisation) and y = J^T W^2 J x on the
    heterogeneous levels 3-5 model against the oracle's restatement of
    ImplicitJtJOperatorFV::Mult (oracle.jtj_matvec: build_perturbation_rhs,
    extract_delta_impedance_complex, build_adjoint_rhs_from_residual and
    compute_sensitivity in the reference's loop), both with inner solves at
    1e-9 so that the multicolour and natural-order SSOR trajectories agree:
    delta Z to 1e-5 of its largest value, y to 1e-4 in norm.  h_cache
    "reference" is the reference's station H of the last background frequency
    for every frequency; "per_frequency" each frequency's own."""
    from oracle.oracle import jtj_matvec
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    freqs = np.array([0.3, 10.0])
    ns = m.count("stations")
    rng = np.random.default_rng(14)
    err = np.full((len(freqs), ns, 4), 0.05) * rng.uniform(0.5, 1.5, (len(freqs), ns, 4))
    err[:, :3, 0] = 0.0                    # unused data
    earth = m.array("sigma") > 1e-6
    x = np.where(earth, rng.standard_normal(m.nc), 0.0)
    gn = GaussNewton(m, freqs, err, inner_tol=1e-9, h_cache=h_cache)
    dz, _ = gn.jvp(x)
    y = gn.matvec(x)
    y_o, dz_o = jtj_matvec(m, freqs, err, x, inner_tol=1e-9, stale_h=(h_cache == "reference"),
                           sigma_bg=wl.sigma_bg, stations=wl.stations[:, :2])
    edz = np.abs(dz - dz_o).max() / np.abs(dz_o).max()
    ey = np.linalg.norm(y - y_o) / np.linalg.norm(y_o)
    print(h_cache, "delta Z", edz, "y", ey, flush=True)
    assert edz <= 1e-5
    assert ey <= 1e-4
    m.close()
