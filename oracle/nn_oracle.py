"""NumPy oracle of the coarse-grid e0 network (forward and gradients).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the device network
(maple3dmt_amd/csrc/device/nn.hip).  It mirrors the device's storage
roundings (BF16 weights, inputs, activations and gradients between layers)
and accumulates in float64.  Its own backward pass is pinned against central
finite differences of the float64 loss in tests/test_nn_oracle.py.

Layout: activations [B, G, G, G, C] (z, y, x), weights per layer
[27, 32, Cin] ([tap][co][ci], tap = 9*(dz+1) + 3*(dy+1) + (dx+1)), biases [32].
"""
from __future__ import annotations

import numpy as np

G = 32
HID = 32
CIN0 = 8
COUT = 12
CINS = (CIN0, HID, HID, HID)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def shift(x, dz, dy, dx):
    """y[v] = x[v + (dz, dy, dx)] with zeros outside (x: [B, G, G, G, C])."""
    out = np.zeros_like(x)
    B, gz, gy, gx, _ = x.shape

    def rng(d, n):
        return (max(0, -d), min(n, n - d)), (max(0, d), min(n, n + d))
    (za, zb), (sa, sb) = rng(dz, gz)
    (ya, yb), (ta, tb) = rng(dy, gy)
    (xa, xb), (ua, ub) = rng(dx, gx)
    if za < zb and ya < yb and xa < xb:
        out[:, za:zb, ya:yb, xa:xb] = x[:, sa:sb, ta:tb, ua:ub]
    return out


def conv(x, w, b, dil):
    out = np.zeros(x.shape[:4] + (w.shape[1],), np.float64)
    for t, dz, dy, dx in taps():
        out += shift(x, dz * dil, dy * dil, dx * dil) @ w[t].T.astype(np.float64)
    return out + b


def conv_grad_input(dz_, w, dil):
    out = np.zeros(dz_.shape[:4] + (w.shape[2],), np.float64)
    for t, dz, dy, dx in taps():
        out += shift(dz_, -dz * dil, -dy * dil, -dx * dil) @ w[t].astype(np.float64)
    return out


def conv_grad_weight(x, dz_, dil):
    cin = x.shape[-1]
    gw = np.zeros((27, dz_.shape[-1], cin), np.float64)
    d2 = dz_.reshape(-1, dz_.shape[-1])
    for t, dz, dy, dx in taps():
        xs = shift(x, dz * dil, dy * dil, dx * dil).reshape(-1, cin)
        gw[t] = d2.T @ xs
    return gw


def unpack(w_flat, b_flat):
    ws, off = [], 0
    for c in CINS:
        n = 27 * HID * c
        ws.append(np.asarray(w_flat[off:off + n], np.float32).reshape(27, HID, c))
        off += n
    bs = [np.asarray(b_flat[l * HID:(l + 1) * HID], np.float64) for l in range(4)]
    return ws, bs


def forward(w_flat, b_flat, x, dil=(1, 2, 4, 1), rounding=True):
    """x: [B, G, G, G, 2] float32 -> (out [B, G, G, G, 12], activations)."""
    ws, bs = unpack(w_flat, b_flat)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    a = np.zeros(x.shape[:4] + (CIN0,), np.float64)
    a[..., :2] = r(x)
    acts = [a]
    for l in range(4):
        z = conv(acts[-1], r(ws[l]), bs[l], dil[l])
        if l < 3:
            acts.append(r(np.maximum(z, 0.0)))
        else:
            acts.append(z[..., :COUT])
    return acts[-1], acts


def gradients(w_flat, b_flat, x, tgt, dil=(1, 2, 4, 1), rounding=True):
    """MSE loss and its gradients in the device's [tap][co][ci] layout."""
    ws, bs = unpack(w_flat, b_flat)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    out, acts = forward(w_flat, b_flat, x, dil, rounding)
    n = out.size
    d = out - tgt
    loss = float(np.mean(d ** 2))
    g = np.zeros(out.shape[:4] + (HID,), np.float64)
    g[..., :COUT] = 2.0 * d / n
    g = r(g)
    gws, gbs = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        dz_ = g
        if l < 3:
            dz_ = g * (acts[l + 1] > 0)
        gws[l] = conv_grad_weight(acts[l], dz_, dil[l])
        gbs[l] = dz_.reshape(-1, HID).sum(axis=0)
        if l > 0:
            g = r(conv_grad_input(dz_, r(ws[l]), dil[l]))
    gw = np.concatenate([gw.ravel() for gw in gws])
    gb = np.concatenate(gbs)
    return loss, gw, gb
