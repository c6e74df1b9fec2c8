"""Station side of the gradient on the device (m3d_impedance_device,
m3d_weighted_residual_device, m3d_adjoint_rhs_device) against the CPU oracle
and the reference formulas, through the C-ABI.

Tolerances: Z and the adjoint right-hand side relative 1e-12 (the same
formulas in the same order; complex division may round differently in the
last bit); the weighted residual and the misfit bit for bit (the same
operations in the same order as inversion_fv.cpp:258-271 and :75-98).
"""
import numpy as np
import pytest

from maple3dmt_amd import configs, inversion
from maple3dmt_amd.native import NativeError
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def block(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    o = Oracle(m, sigma_bg=wl.sigma_bg)
    o.set_sigma(m.array("sigma"))
    o.find_stations(wl.stations[:, :2])
    return wl, m, o


def _cvec(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_device_station_steps_match_oracle(block):
    import torch
    dev = torch.device("cuda:0")
    wl, m, o = block
    rng = np.random.default_rng(21)
    freqs = np.array([1e-3, 0.3, 10.0])
    nf, ns = len(freqs), len(wl.stations)
    E = _cvec(rng, 2 * nf, m.n)
    obs = _cvec(rng, nf, ns, 4) * 0.05
    err = np.abs(rng.standard_normal((nf, ns, 4))) * 0.01 + 1e-3
    err[0, 3, 1] = 0.0          # unused datum (error <= 0)
    wgt = np.ones((nf, ns, 4))
    wgt[1, 5, 2] = 0.0          # unused datum (weight <= 0)
    st = m.device_stations(4)
    d_e = torch.from_numpy(E).to(dev)
    d_z = torch.zeros((nf, ns, 4), dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    st.impedance(2 * np.pi * freqs, d_e.data_ptr(), d_z.data_ptr())
    Z = d_z.cpu().numpy()
    d_obs = torch.from_numpy(obs).to(dev)
    d_err = torch.from_numpy(err).to(dev)
    d_w = torch.from_numpy(wgt).to(dev)
    d_wr = torch.zeros_like(d_z)
    d_phi = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    st.weighted_residual(nf, d_obs.data_ptr(), d_err.data_ptr(), d_w.data_ptr(), d_wr.data_ptr(),
                         d_phi.data_ptr())
    d_rhs = torch.full((2 * nf, m.n), 7.0 + 0j, dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    st.adjoint_rhs(nf, d_wr.data_ptr(), d_rhs.data_ptr())
    wr, rhs, phi = d_wr.cpu().numpy(), d_rhs.cpu().numpy(), float(d_phi.cpu().item())
    # Weighted residual and misfit: the reference's operations in order.
    phi_ref = 0.0
    for f in range(nf):
        for s in range(ns):
            for c in range(4):
                if wgt[f, s, c] <= 0 or err[f, s, c] <= 0:
                    assert wr[f, s, c] == 0
                    continue
                r = obs[f, s, c] - Z[f, s, c]
                w = 1.0 / err[f, s, c]
                assert wr[f, s, c] == w * w * np.conj(r)
                phi_ref += 0.5 * w * w * (r.real * r.real + r.imag * r.imag)
    assert phi == phi_ref
    for f, fr in enumerate(freqs):
        w = 2 * np.pi * fr
        o.assemble(w)
        zo = o.impedance(E[2 * f], E[2 * f + 1])
        assert np.abs(Z[f] - zo).max() <= 1e-12 * np.abs(zo).max(), f
        q1, q2 = o.adjoint_rhs(wr[f])
        assert np.abs(rhs[2 * f] - q1).max() <= 1e-12 * np.abs(q1).max(), f
        assert np.abs(rhs[2 * f + 1] - q2).max() <= 1e-12 * np.abs(q2).max(), f
        off = np.ones(m.n, bool)
        off[m.array("st_uedge")] = False
        assert not rhs[2 * f][off].any() and not rhs[2 * f + 1][off].any()
    # Calls out of order or beyond the set's size are refused.
    with pytest.raises(NativeError):
        st.impedance(2 * np.pi * np.ones(5), d_e.data_ptr(), d_z.data_ptr())
    with pytest.raises(NativeError):
        st.adjoint_rhs(nf + 1, d_wr.data_ptr(), d_rhs.data_ptr())


def test_device_resident_gradient_matches_oracle_loop(block):
    """The gradient with the station steps on the device against the
    oracle's per-frequency loop (inversion_fv.cpp:250-290), and against the
    host-station path of the product."""
    wl, m, o = block
    freqs = wl.freqs
    zs = [o.solve_frequency(f)[2] for f in freqs]
    obs = np.stack(zs) * (1.05 + 0.02j)
    floor = 0.05 * np.maximum(np.abs(obs[:, :, 1]), np.abs(obs[:, :, 2]))
    err = np.repeat(floor[:, :, None], 4, axis=2)
    res = inversion.gradient_device(m, freqs, obs, err)
    host = inversion.gradient(m, freqs, obs, err)
    phi_o, g_o = 0.0, np.zeros(m.nc)
    for fi, f in enumerate(freqs):
        e1, e2, z, _ = o.solve_frequency(f)
        r = obs[fi] - z
        phi_o += 0.5 * float(np.sum(np.abs(r) ** 2 / err[fi] ** 2))
        q1, q2 = o.adjoint_rhs(np.conj(r) / err[fi] ** 2)
        g_o += o.sensitivity(e1, o.adjoint(q1)[0]) + o.sensitivity(e2, o.adjoint(q2)[0])
    assert abs(res.phi - phi_o) <= 2e-3 * phi_o
    assert np.linalg.norm(res.grad - g_o) <= 2e-3 * np.linalg.norm(g_o)
    assert abs(res.phi - host.phi) <= 1e-12 * host.phi
    # Adjoint trajectories from right-hand sides equal to 1e-12 agree to the
    # solve tolerance (test_gpu_gradient.py).
    assert np.linalg.norm(res.grad - host.grad) <= 2e-3 * np.linalg.norm(host.grad)
    assert np.abs(res.z_pred - host.z_pred).max() <= 1e-12 * np.abs(host.z_pred).max()
