"""Device network (BF16 MFMA conv3d) against the NumPy oracle, device e0
prolongation against NumPy, and the initial-guess path of the FORWARD solve.

Tolerances: the device and the oracle apply the same BF16 roundings to
weights, inputs, activations and inter-layer gradients, but accumulate in FP32
(device) vs float64 (oracle); a BF16 rounding that flips near a tie moves a
value by 2^-8 relative, so outputs are compared to 1e-2 and gradients to 3e-2
in relative norm.
"""
import numpy as np
import pytest

from maple3dmt_amd import configs
from maple3dmt_amd.learned import Box, CoarseNet, deposit, initial_guess
from maple3dmt_amd.native import MODE_FORWARD, SolveParams
from oracle import nn_oracle as no
from conftest import solved

pytestmark = pytest.mark.gpu

G = 32


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


def _random_params(net, rng):
    w, b = net.params()
    ws, _ = no.unpack(w, b)
    for l, x in enumerate(ws):
        x[...] = rng.standard_normal(x.shape).astype(np.float32) * (0.6 / np.sqrt(27 * x.shape[2]))
    ws[0][:, :, 2:] = 0
    ws[3][:, 12:, :] = 0
    w = np.concatenate([x.ravel() for x in ws])
    b = (rng.standard_normal(b.shape) * 0.05).astype(np.float32)
    net.set_params(w, b)
    return w, b


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_and_gradients_match_oracle(ctx, torch_dev):
    torch, dev = torch_dev
    rng = np.random.default_rng(0)
    net = CoarseNet(ctx, max_batch=2)
    w, b = _random_params(net, rng)
    x = rng.standard_normal((2, G, G, G, 2)).astype(np.float32)
    t = (rng.standard_normal((2, G, G, G, 12)) * 0.5).astype(np.float32)
    dx = torch.from_numpy(x.reshape(2, -1, 2)).to(dev)
    dt = torch.from_numpy(t.reshape(2, -1, 12)).to(dev)
    out = torch.empty((2, G ** 3, 12), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    net.forward(dx.data_ptr(), out.data_ptr(), 2)
    ref, _ = no.forward(w, b, x, dil=net.dil)
    assert _rel(out.cpu().numpy().reshape(ref.shape), ref) <= 1e-2
    loss = net.gradients(dx.data_ptr(), dt.data_ptr(), 2)
    gw, gb = net.params(grads=True)
    rloss, rgw, rgb = no.gradients(w, b, x, t, dil=net.dil)
    assert abs(loss - rloss) <= 1e-3 * rloss
    offs = np.cumsum([0] + [27 * 32 * c for c in no.CINS])
    for l in range(4):
        seg = slice(offs[l], offs[l + 1])
        assert _rel(gw[seg], rgw[seg]) <= 3e-2, (l, _rel(gw[seg], rgw[seg]))
        assert _rel(gb[l * 32:(l + 1) * 32], rgb[l * 32:(l + 1) * 32]) <= 3e-2
    net.close()


def test_adam_step_matches_update_rule(ctx, torch_dev):
    """First Adam step from zero moments: m_hat = g, v_hat = g^2, so every
    parameter moves by -lr * g / (|g| + eps) with g the device gradient."""
    torch, dev = torch_dev
    rng = np.random.default_rng(4)
    net = CoarseNet(ctx, max_batch=1, seed=5)
    w0, b0 = net.params()
    x = rng.standard_normal((1, G ** 3, 2)).astype(np.float32)
    t = (rng.standard_normal((1, G ** 3, 12)) * 0.3).astype(np.float32)
    dx = torch.from_numpy(x).to(dev)
    dt = torch.from_numpy(t).to(dev)
    torch.cuda.synchronize()
    net.gradients(dx.data_ptr(), dt.data_ptr(), 1)
    gw, gb = net.params(grads=True)
    lr = 1e-3
    net.train_step(dx.data_ptr(), dt.data_ptr(), 1, lr)
    w1, b1 = net.params()
    ew = w0 - lr * gw / (np.abs(gw) + 1e-8)
    eb = b0 - lr * gb / (np.abs(gb) + 1e-8)
    assert np.abs(w1 - ew).max() <= 1e-6
    assert np.abs(b1 - eb).max() <= 1e-6
    net.close()


def test_training_reduces_loss(ctx, torch_dev):
    torch, dev = torch_dev
    rng = np.random.default_rng(1)
    net = CoarseNet(ctx, max_batch=2, seed=3)
    # Smooth inputs (sums of a few low wavenumber modes) and a target that is
    # a smooth function of them, learnable by a few conv layers.
    z, y, xg = np.meshgrid(*(np.arange(G) / G,) * 3, indexing="ij")
    x = np.zeros((2, G, G, G, 2), np.float32)
    for b in range(2):
        for c in range(2):
            k = rng.integers(1, 3, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            x[b, ..., c] = (np.sin(2 * np.pi * k[0] * z + ph[0]) * np.cos(2 * np.pi * k[1] * y + ph[1])
                            + 0.5 * np.sin(2 * np.pi * k[2] * xg + ph[2]))
    t = np.zeros((2, G, G, G, 12), np.float32)
    t[..., 0] = 0.5 * x[..., 0]
    t[..., 7] = -0.25 * x[..., 1] + 0.1
    t[..., 3] = 0.2 * np.maximum(x[..., 0], 0)
    dx = torch.from_numpy(x.reshape(2, -1, 2)).to(dev)
    dt = torch.from_numpy(t.reshape(2, -1, 12)).to(dev)
    torch.cuda.synchronize()
    losses = [net.train_step(dx.data_ptr(), dt.data_ptr(), 2, 2e-3) for _ in range(150)]
    assert min(losses[-10:]) < 0.5 * losses[0], (losses[0], losses[-10:])
    net.close()


def test_prolongation_matches_numpy(ctx, torch_dev):
    torch, dev = torch_dev
    wl = configs.halfspace()
    m = configs.build_model(wl, ctx)
    box = Box()
    rng = np.random.default_rng(2)
    field = rng.standard_normal((1, G ** 3, 12)).astype(np.float32)
    net = CoarseNet(ctx, max_batch=1)
    exyz = np.stack([m.array("edge_x"), m.array("edge_y"), m.array("edge_z")], 1).astype(np.float32)
    edir = m.array("edge_dir")
    d_f = torch.from_numpy(field).to(dev)
    d_xyz = torch.from_numpy(exyz).to(dev)
    d_dir = torch.from_numpy(edir).to(dev)
    x0 = torch.zeros((2, m.n), dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    net.prolong(d_f.data_ptr(), 1, m.n, d_xyz.data_ptr(), d_dir.data_ptr(), box, x0.data_ptr())
    got = x0.cpu().numpy()
    h = (box.x1 - box.x0) / G, (box.y1 - box.y0) / G, (box.z1 - box.z0) / G
    f = field[0].reshape(G, G, G, 12)
    ref = np.zeros((2, m.n), np.complex128)
    for e in range(m.n):
        fx = (exyz[e, 0] - box.x0) / h[0] - 0.5
        fy = (exyz[e, 1] - box.y0) / h[1] - 0.5
        fz = (exyz[e, 2] - box.z0) / h[2] - 0.5
        if min(fx, fy, fz) < 0 or max(fx, fy, fz) > G - 1:
            continue
        ix, iy, iz = min(int(fx), G - 2), min(int(fy), G - 2), min(int(fz), G - 2)
        tx, ty, tz = fx - ix, fy - iy, fz - iz
        v = np.zeros(12)
        for c in range(8):
            cx, cy, cz = c & 1, (c >> 1) & 1, c >> 2
            wgt = (tx if cx else 1 - tx) * (ty if cy else 1 - ty) * (tz if cz else 1 - tz)
            v += wgt * f[iz + cz, iy + cy, ix + cx]
        d = edir[e]
        ref[0, e] = v[2 * d] + 1j * v[2 * d + 1]
        ref[1, e] = v[6 + 2 * d] + 1j * v[6 + 2 * d + 1]
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    net.close()


def test_forward_initial_guess_and_fallback(ctx, torch_dev):
    torch, dev = torch_dev
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3,
                          station_radius=5e3)
    m = configs.build_model(wl, ctx)
    omegas = np.repeat(2 * np.pi * wl.freqs, 2)
    rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
    aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(omegas)])
    prm = m.fwd_params.solve_params()
    x_ref, r_ref = m.device.solve(MODE_FORWARD, omegas, rhs, aux, params=prm)
    esec = x_ref - aux
    # The converged E_sec as the guess: kept, and far fewer iterations.
    prm_g = m.fwd_params.solve_params()
    prm_g.use_initial_guess = 1
    x1, r1 = m.device.solve(MODE_FORWARD, omegas, rhs, aux, x=esec, params=prm_g)
    assert all(r["guess_used"] and solved(r) for r in r1)
    assert sum(r["iterations"] for r in r1) < 0.2 * sum(r["iterations"] for r in r_ref)
    assert np.abs(x1 - x_ref).max() <= 1e-4 * np.abs(x_ref).max()
    # A useless guess falls back to the zero start and gives the same answer.
    bad = esec * 1e6
    x2, r2 = m.device.solve(MODE_FORWARD, omegas, rhs, aux, x=bad, params=prm_g)
    assert not any(r["guess_used"] for r in r2)
    assert [r["iterations"] for r in r2] == [r["iterations"] for r in r_ref]
    # Least-squares scaling (mode 2) recovers the mis-scaled guess ...
    prm_s = m.fwd_params.solve_params()
    prm_s.use_initial_guess = 2
    x3, r3 = m.device.solve(MODE_FORWARD, omegas, rhs, aux, x=bad * (0.3 - 0.2j), params=prm_s)
    assert all(r["guess_used"] and solved(r) for r in r3)
    assert all(r["guess_residual"] < 1e-5 for r in r3)
    assert sum(r["iterations"] for r in r3) < 0.2 * sum(r["iterations"] for r in r_ref)
    assert np.abs(x3 - x_ref).max() <= 1e-4 * np.abs(x_ref).max()
    # ... and rejects noise, which cannot halve the residual.
    noise = np.random.default_rng(3).standard_normal(esec.shape) * np.abs(esec).max()
    x4, r4 = m.device.solve(MODE_FORWARD, omegas, rhs, aux, x=noise.astype(np.complex128), params=prm_s)
    assert not any(r["guess_used"] for r in r4)
    assert all(r["guess_residual"] > 0.5 for r in r4)
    assert [r["iterations"] for r in r4] == [r["iterations"] for r in r_ref]


def test_trained_network_on_held_out_models(ctx, torch_dev):
    """The committed network (trained on solver-generated fields of seeded
    random models, levels 3-5 to 3-7, whole-domain voxel box) on held-out
    seeds: the fast path from the zero start and from the network's e0 at the
    same tolerance.  Both converge, their impedances agree (tol 1e-9, so the
    two solutions differ by less than 1e-6), the fallback keeps the guess only
    where its least-squares scaled residual is below 0.5 of ||b|| and starts
    from zero elsewhere; iterations and solves/s of both starts are printed."""
    import os
    import time
    from maple3dmt_amd.learned import domain_box
    from maple3dmt_amd.train_e0 import EVAL_SEED0
    torch, dev = torch_dev
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "maple3dmt_amd", "data",
                        "e0_net.npz")
    d = np.load(path)
    net = CoarseNet(ctx, max_batch=8, dilations=tuple(int(v) for v in d["dilations"]))
    net.set_params(d["w"], d["b"])
    box = Box(*d["box"].tolist())
    tot = {"zero": [0, 0.0], "e0": [0, 0.0]}
    for seed, level in ((EVAL_SEED0 + 150, 5), (EVAL_SEED0 + 151, 6), (EVAL_SEED0 + 152, 6)):
        wl = configs.random3d(seed, max_level=level, half_width=100e3, station_radius=20e3, maxiter=40000)
        assert np.allclose(box.arr(), domain_box(wl).arr())
        m = configs.build_model(wl, ctx)
        om = np.repeat(2 * np.pi * wl.freqs, 2)
        rhs = np.stack([m.rhs_total(m.bc_values(w, p % 2)) for p, w in enumerate(om)])
        aux = np.stack([m.primary_field(w, p % 2) for p, w in enumerate(om)])
        d_rhs, d_aux = torch.from_numpy(rhs).to(dev), torch.from_numpy(aux).to(dev)
        zs = {}
        for start in ("zero", "e0"):
            prm = wl.fwd.fast_params(tol=1e-9)
            t = time.perf_counter()
            if start == "zero":
                d_x = torch.zeros_like(d_rhs)
            else:
                d_x = initial_guess(net, m, wl, box)
                prm.use_initial_guess = 2
            torch.cuda.synchronize()
            res = m.device.solve_device(MODE_FORWARD, om, d_rhs.data_ptr(), d_aux.data_ptr(), d_x.data_ptr(),
                                        len(om), prm)
            dt = time.perf_counter() - t
            assert all(solved(r, 1e-9) for r in res), res
            if start == "e0":
                for r in res:
                    assert r["guess_used"] == (r["guess_residual"] < 0.5), r
            x = d_x.cpu().numpy()
            zs[start] = np.stack([m.extract_impedance(om[2 * f], x[2 * f], x[2 * f + 1])
                                  for f in range(len(wl.freqs))])
            its = [r["iterations"] for r in res]
            tot[start][0] += sum(its)
            tot[start][1] += dt
            print(f"seed {seed} level {level} {m.n} edges {start}: {sum(its)} iterations, {dt:.3f}s"
                  + (f", guess used {sum(r['guess_used'] for r in res)}/16" if start == "e0" else ""), flush=True)
        rel = np.abs(zs["e0"][..., 1:3] - zs["zero"][..., 1:3]) / np.abs(zs["zero"][..., 1:3])
        assert rel.max() <= 1e-6, rel.max()
        m.close()
    for k, (it, sec) in tot.items():
        print(f"{k} start: {it} iterations, {48 / sec:.1f} solves/s", flush=True)
    net.close()
