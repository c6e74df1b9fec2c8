"""The allocation-free sensitivity entry point (m3d_sensitivity_device_noalloc)
against m3d_sensitivity_device through the C-ABI: two calls in a row, the
second with fewer systems, each equal bit for bit to the allocating entry
point (the same kernel reads the same frequencies, staged elsewhere); sizes
beyond its scratch are refused."""
import numpy as np
import pytest

from maple3dmt_amd import configs
from maple3dmt_amd.native import Context, NativeError

pytestmark = pytest.mark.gpu


def test_noalloc_sensitivity_matches_allocating_entry_point():
    import torch
    dev = torch.device("cuda:0")
    ctx = Context(0)
    wl = configs.random3d(8, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    m = configs.build_model(wl, ctx)
    rng = np.random.default_rng(17)
    for nsys in (6, 3):
        print("nsys", nsys, flush=True)
        e = rng.standard_normal((nsys, m.n)) + 1j * rng.standard_normal((nsys, m.n))
        lam = rng.standard_normal((nsys, m.n)) + 1j * rng.standard_normal((nsys, m.n))
        om = 2 * np.pi * np.logspace(-3, 2, nsys)
        d_e = torch.from_numpy(e).to(dev)
        d_l = torch.from_numpy(lam).to(dev)
        d_a = torch.full((m.nc,), np.nan, dtype=torch.float64, device=dev)
        d_b = torch.full((m.nc,), np.nan, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        m.device.sensitivity_device(om, d_e.data_ptr(), d_l.data_ptr(), d_a.data_ptr())
        m.device.sensitivity_device_noalloc(om, d_e.data_ptr(), d_l.data_ptr(), d_b.data_ptr())
        a, b = d_a.cpu().numpy(), d_b.cpu().numpy()
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        assert np.array_equal(a, b)
        ref = sum(m.sensitivity(om[k], e[k], lam[k]) for k in range(nsys))
        assert np.linalg.norm(b - ref) <= 1e-12 * np.linalg.norm(ref)
    with pytest.raises(NativeError, match="exceed its scratch"):
        m.device.sensitivity_device_noalloc(np.ones(4097), d_e.data_ptr(), d_l.data_ptr(), d_b.data_ptr())
    m.close()
    ctx.close()
