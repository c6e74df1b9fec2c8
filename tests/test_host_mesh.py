"""Host producer invariants: octree, staggered grid, discrete operators.

The reference's mesh comes from p4est, which is absent here, so the mesh is
pinned by what the mathematics fixes: counts on uniform meshes, Morton order,
2:1 face balance, div(curl) = 0 (examples/test_octree_basic.cpp:101-134),
curl(grad) = 0 (operators.cpp:846-847), the closed-form uniform-cube stencil
(S_ii = 4h/mu0, S_ij = +-h/mu0, Me = sigma h^3) and symmetry.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from maple3dmt_amd import configs
from maple3dmt_amd.native import ForwardParams, MeshParams, Model

MU0 = 4e-7 * np.pi


def csr(m, name, shape):
    return sp.csr_matrix((m.array(name + "_val"), m.array(name + "_col"), m.array(name + "_ptr")),
                         shape=shape)


@pytest.fixture(scope="module")
def uniform():
    mp = MeshParams(domain=(-4e3, 4e3, -4e3, 4e3, -4e3, 4e3), min_level=3, max_level=3)
    m = Model(mp, np.zeros((0, 3)), [1.0])
    m.set_sigma(np.full(m.nc, 0.01))
    return m


@pytest.fixture(scope="module")
def adaptive():
    wl = configs.halfspace()
    return configs.build_model(wl)


def test_uniform_counts(uniform):
    n = 8
    assert uniform.nc == n ** 3
    assert uniform.n == 3 * n * (n + 1) ** 2
    assert uniform.count("faces") == 3 * n ** 2 * (n + 1)
    assert uniform.nn == (n + 1) ** 3
    assert uniform.count("hanging_faces") == 0


def test_morton_order(uniform):
    # First cell at the (x_min, y_min, z_min) corner, then +x, +y, +z.
    cx, cy, cz = (uniform.array(k) for k in ("cell_cx", "cell_cy", "cell_cz"))
    h = 1000.0
    assert (cx[0], cy[0], cz[0]) == (-3500.0, -3500.0, -3500.0)
    assert (cx[1] - cx[0], cy[2] - cy[0], cz[4] - cz[0]) == (h, h, h)
    assert (cx[8], cy[8], cz[8]) == (-3500.0 + 2 * h, -3500.0, -3500.0)


def test_face_balance_and_hanging(adaptive):
    m = adaptive
    lev = m.array("cell_level")
    fm, fp = m.array("face_minus"), m.array("face_plus")
    both = (fm >= 0) & (fp >= 0)
    dl = np.abs(lev[fm[both]] - lev[fp[both]])
    assert dl.max() <= 1
    hf = m.array("hanging_faces").reshape(-1, 5)
    assert len(hf) > 0
    dead_face = m.array("dead_face")
    assert dead_face[hf[:, 0]].all()
    # Every hanging edge pair: children parallel, half length, collinear.
    he = m.array("hanging_edges").reshape(-1, 3)
    L = m.array("edge_len")
    d = m.array("edge_dir")
    xyz = np.stack([m.array("edge_x"), m.array("edge_y"), m.array("edge_z")], 1)
    assert (d[he[:, 1]] == d[he[:, 0]]).all() and (d[he[:, 2]] == d[he[:, 0]]).all()
    assert np.allclose(L[he[:, 1]], L[he[:, 0]] / 2) and np.allclose(L[he[:, 2]], L[he[:, 0]] / 2)
    assert np.allclose(0.5 * (xyz[he[:, 1]] + xyz[he[:, 2]]), xyz[he[:, 0]])


def test_div_curl_zero(uniform, adaptive):
    for m in (uniform, adaptive):
        Ce = csr(m, "Ce", (m.count("faces"), m.n))
        D = csr(m, "D", (m.nc, m.count("faces")))
        e = np.sin(np.arange(m.n) * 0.1) + np.cos(np.arange(m.n) * 0.03)
        # Ce carries +-L entries: divide per edge to get the topological curl.
        Ct = Ce @ sp.diags(1.0 / m.array("edge_len"))
        # Holds on the adaptive mesh too (examples/test_octree_basic.cpp:101-134
        # asserts it there): hanging faces sum their children's circulations.
        assert np.linalg.norm(D @ (Ct @ e)) <= 1e-12 * np.linalg.norm(e)
        assert np.linalg.norm(D @ (Ce @ e)) <= 1e-14 * abs(Ce).max() * np.linalg.norm(e)


def test_curl_grad_zero(uniform, adaptive):
    for m in (uniform, adaptive):
        Ce = csr(m, "Ce", (m.count("faces"), m.n))
        G = sp.csr_matrix((m.array("G_val"), m.array("G_col"), m.array("G_ptr")), shape=(m.n, m.nn))
        phi = np.cos(np.arange(m.nn) * 0.37)
        g = G @ phi
        # Dead edges are interpolated from their children, so the identity
        # holds for the rewritten curl on every alive face.
        assert np.linalg.norm(Ce @ g) <= 1e-10 * np.linalg.norm(Ce.data) * np.linalg.norm(g)


def test_uniform_stencil_closed_form(uniform):
    m = uniform
    h = 1000.0
    S = csr(m, "S", (m.n, m.n))
    bc = m.array("is_bc").astype(bool)
    interior = np.where(~bc)[0]
    d = S.diagonal()
    assert np.allclose(d[interior], 4 * h / MU0, rtol=1e-12)
    off = S - sp.diags(d)
    rows = off[interior]
    assert np.allclose(np.abs(rows.data), h / MU0, rtol=1e-12)
    assert (np.diff(rows.indptr) == 12).all()
    me = m.array("me")
    assert np.allclose(me[interior], 0.01 * h ** 3, rtol=1e-12)


def test_stiffness_symmetric(adaptive):
    m = adaptive
    S = csr(m, "S", (m.n, m.n))
    diff = abs(S - S.T)
    assert diff.max() <= 1e-14 * abs(S).max()


def test_offdiag_is_free_free_restriction(adaptive):
    m = adaptive
    S = csr(m, "S", (m.n, m.n)).tolil()
    cons = m.array("is_constrained").astype(bool)
    off = csr(m, "offdiag", (m.n, m.n))
    free = np.where(~cons)[0]
    Sf = csr(m, "S", (m.n, m.n))[free][:, free]
    Sf = Sf - sp.diags(Sf.diagonal())
    Of = off[free][:, free]
    assert abs(Sf - Of).max() == 0.0
    assert off[np.where(cons)[0]].nnz == 0
    assert np.array_equal(m.array("diag_s")[free], csr(m, "S", (m.n, m.n)).diagonal()[free])


def test_primary_field_closed_form(adaptive):
    m = adaptive
    w = 2 * np.pi * 0.7
    e = m.primary_field(w, 0)
    a = np.sqrt(1j * w * MU0 * 0.01)
    z = m.array("edge_z")
    d = m.array("edge_dir")
    ref = np.where(z >= 0, 1 + a * z, np.exp(a * np.minimum(z, 0)))
    assert np.allclose(e[d == 0], ref[d == 0], rtol=1e-14, atol=0)
    assert (e[d != 0] == 0).all()
    assert a.real > 0  # decays into the earth


def test_flags(adaptive):
    m = adaptive
    z = m.array("edge_z")
    bc = m.array("is_bc").astype(bool)
    assert bc[z > 1.0].all()
    dc = m.array("dc_bc_edge").astype(bool)
    dead = m.array("is_dead").astype(bool)
    assert not (dc & dead).any()
    assert (dc | dead)[bc].all()


def test_random_workload_mesh_is_balanced():
    wl = configs.random3d(3, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                          station_radius=5e3)
    m = configs.build_model(wl)
    lev = m.array("cell_level")
    fm, fp = m.array("face_minus"), m.array("face_plus")
    both = (fm >= 0) & (fp >= 0)
    assert np.abs(lev[fm[both]] - lev[fp[both]]).max() <= 1
    s = m.array("sigma")
    assert s.min() == pytest.approx(1e-8) and len(np.unique(s)) >= 3


def test_compact_station_steps_equal_full_field_forms():
    """Z and the adjoint right-hand sides from the values on the station
    cells' edges only (the device-resident gradient's path) equal the
    full-field forms bit for bit (extract_impedance_, :1490-1727;
    build_adjoint_rhs_from_residual, :1804-2013)."""
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                          station_radius=5e3)
    m = configs.build_model(wl)
    idx = m.station_edges()
    assert len(idx) > 0 and np.all(np.diff(idx) > 0)
    rng = np.random.default_rng(4)
    e1 = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    e2 = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    w = 2 * np.pi * 0.7
    z_full = m.extract_impedance(w, e1, e2)
    wr = rng.standard_normal(z_full.shape) + 1j * rng.standard_normal(z_full.shape)
    r1, r2 = m.adjoint_rhs(w, wr)
    z_c = m.extract_impedance_compact(w, e1[idx], e2[idx])
    c1, c2 = m.adjoint_rhs_compact(w, wr)
    assert np.array_equal(z_c, z_full)
    assert np.array_equal(c1, r1[idx]) and np.array_equal(c2, r2[idx])
    rest = np.ones(m.n, bool)
    rest[idx] = False
    assert not np.any(r1[rest]) and not np.any(r2[rest])


def test_concurrent_builds_equal_sequential():
    """Host models built on several threads at once (multimodel.build_models)
    hold exactly the arrays of the same models built one after another."""
    from maple3dmt_amd import multimodel
    wls = [configs.random3d(20 + i, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                            station_radius=5e3) for i in range(4)]
    seq = [configs.build_model(wl, None) for wl in wls]
    par = multimodel.build_models(wls, [None], n_threads=4)
    names = ("offdiag_ptr", "offdiag_col", "offdiag_val", "diag_s", "me", "sigma", "cell_cx", "L_val",
             "G_col", "is_bc")
    for a, b in zip(seq, par):
        assert a.n == b.n and a.nc == b.nc
        for name in names:
            assert np.array_equal(a.array(name), b.array(name)), name
