"""Sigma update on the device (a line-search trial) against the host update.

m3d_operator_update_sigma recomputes Me, the DivCorr weights and the values
of the DivCorr Laplacian on the device in the host builders' order, so every
device array equals the host path's bit for bit.  Checked through what the
arrays feed: forward solves (Me in the operator, the Laplacian in the
post-hoc DivCorr), m3d_divcorr, and the device-resident gradient (the
sensitivity's sigma) must be bitwise identical after either update.
"""
import numpy as np
import pytest

from maple3dmt_amd import configs, inversion
from maple3dmt_amd.multimodel import halfspace_data
from maple3dmt_amd.native import NativeError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(ctx):
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                          station_radius=5e3)
    return wl, configs.build_model(wl, ctx)


def _trial_sigma(m, seed):
    rng = np.random.default_rng(seed)
    s = m.array("sigma").copy()
    earth = s > 1e-6
    s[earth] *= np.exp(rng.uniform(-0.7, 0.7, earth.sum()))
    return s


def _solve(m, wl):
    E, Z, res = m.solve_frequencies(wl.freqs)
    return E, Z, [r["iterations"] for r in res]


def test_device_update_matches_host_update(model):
    wl, m = model
    s0 = m.array("sigma").copy()
    s1 = _trial_sigma(m, 1)
    rng = np.random.default_rng(2)
    e = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    om = [2 * np.pi * wl.freqs[0]] * 2
    try:
        m.set_sigma(s1)
        E_h, Z_h, it_h = _solve(m, wl)
        dc_h = m.device.divcorr(om, e)
        m.set_sigma(s0)
        _solve(m, wl)                       # a different state in between
        m.update_sigma_device(s1)
        assert np.array_equal(m.array("sigma"), s1)
        E_d, Z_d, it_d = _solve(m, wl)
        dc_d = m.device.divcorr(om, e)
        assert it_d == it_h
        assert np.array_equal(E_d, E_h)
        assert np.array_equal(Z_d, Z_h)
        assert np.array_equal(dc_d, dc_h)
    finally:
        m.set_sigma(s0)


def test_device_update_gradient_matches_host_update(model):
    wl, m = model
    s0 = m.array("sigma").copy()
    s2 = _trial_sigma(m, 3)
    obs, err = halfspace_data(wl)
    try:
        m.set_sigma(s2)
        g_h = inversion.gradient_device(m, wl.freqs, obs, err)
        m.set_sigma(s0)
        m.update_sigma_device(s2)
        g_d = inversion.gradient_device(m, wl.freqs, obs, err)
        assert g_d.phi == g_h.phi
        assert np.array_equal(g_d.grad, g_h.grad)
    finally:
        m.set_sigma(s0)


def test_pattern_change_is_refused(model):
    wl, m = model
    s0 = m.array("sigma").copy()
    try:
        # Conductivities so small that every DivCorr edge weight falls below
        # the 1e-30 threshold that decided the Laplacian's pattern.
        with pytest.raises(NativeError, match="pattern"):
            m.update_sigma_device(np.full_like(s0, 1e-60))
    finally:
        m.set_sigma(s0)
    # The host update restores a working operator.
    _, _, res = m.solve_frequencies(wl.freqs[-1:])
    assert all(r["converged"] for r in res)
