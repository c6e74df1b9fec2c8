"""Pins the NumPy network oracle: its backward pass against central finite
differences of its own float64 loss (no BF16 rounding), and its conv against
a direct loop on a tiny grid."""
import numpy as np

from oracle import nn_oracle as no


def _params(rng, scale=0.2):
    nw = 27 * 32 * (8 + 3 * 32)
    w = (rng.standard_normal(nw) * scale).astype(np.float32)
    b = (rng.standard_normal(4 * 32) * 0.1).astype(np.float32)
    # Padded input channels 2..7 and output channels 12..31 are unused.
    ws, _ = no.unpack(w, b)
    ws[0][:, :, 2:] = 0
    ws[3][:, 12:, :] = 0
    return np.concatenate([x.ravel() for x in ws]), b


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 5, 5, 5, 3))
    w = rng.standard_normal((27, 4, 3)).astype(np.float32)
    b = rng.standard_normal(4)
    out = no.conv(x, w, b, dil=2)
    ref = np.zeros((1, 5, 5, 5, 4))
    for z in range(5):
        for y in range(5):
            for xx in range(5):
                acc = b.copy()
                for t, dz, dy, dx in no.taps():
                    zz, yy, x2 = z + 2 * dz, y + 2 * dy, xx + 2 * dx
                    if 0 <= zz < 5 and 0 <= yy < 5 and 0 <= x2 < 5:
                        acc += w[t].astype(np.float64) @ x[0, zz, yy, x2]
                ref[0, z, y, xx] = acc
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    w, b = _params(rng)
    x = rng.standard_normal((1, 6, 6, 6, 2)).astype(np.float32)
    tgt = rng.standard_normal((1, 6, 6, 6, 12))
    dil = (1, 2, 1, 1)
    loss, gw, gb = no.gradients(w, b, x, tgt, dil=dil, rounding=False)
    eps = 1e-3
    wd = w.astype(np.float64)
    offs = np.cumsum([0] + [27 * 32 * c for c in no.CINS])
    picks = []
    for l in range(4):
        seg = np.arange(offs[l], offs[l + 1])
        nz = seg[np.abs(gw[seg]) > 1e-6 * np.abs(gw).max()]
        picks += list(rng.choice(nz, 3, replace=False))
    for i in picks:
        wp = wd.copy(); wp[i] += eps
        wm = wd.copy(); wm[i] -= eps
        lp = np.mean((no.forward(wp, b, x, dil, False)[0] - tgt) ** 2)
        lm = np.mean((no.forward(wm, b, x, dil, False)[0] - tgt) ** 2)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - gw[i]) <= 1e-4 * max(abs(fd), 1e-8) + 1e-9, (i, fd, gw[i])
    for i in (0, 33, 70, 99):
        bp = b.astype(np.float64).copy(); bp[i] += eps
        bm = b.astype(np.float64).copy(); bm[i] -= eps
        fd = (np.mean((no.forward(wd, bp, x, dil, False)[0] - tgt) ** 2) -
              np.mean((no.forward(wd, bm, x, dil, False)[0] - tgt) ** 2)) / (2 * eps)
        assert abs(fd - gb[i]) <= 1e-4 * max(abs(fd), 1e-8) + 1e-9


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159], np.float32)
    r = no.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] + 3.140625) < 1e-7
