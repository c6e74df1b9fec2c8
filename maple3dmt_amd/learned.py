"""Learned initial field e0 for the forward solve.

A coarse-grid neural operator (include/m3d_net.h) maps the conductivity,
resampled onto a 32^3 box around the station array, and log10(f) to the
scattered field of both polarisations at the box's voxel centres.  Trilinear
interpolation to edge midpoints gives e0, which the FORWARD solve uses as its
starting E_sec wherever ||b - A e0|| < ||b|| (else the reference's zero start).

Training data are the solver's own forward solutions: E_sec = E - E_pri,
deposited onto the voxel centres with trilinear weights.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import configs
from .native import Context, Model, NativeError, lib, _ptr

G = 32
NCH = 12


@dataclass
class Box:
    """Voxel box of the network; voxel centres at x0 + (i + 1/2) hx."""
    x0: float = -40e3
    x1: float = 40e3
    y0: float = -40e3
    y1: float = 40e3
    z0: float = -56e3
    z1: float = 24e3

    def arr(self) -> np.ndarray:
        return np.array([self.x0, self.x1, self.y0, self.y1, self.z0, self.z1], np.float64)

    def centres(self):
        hx, hy, hz = (self.x1 - self.x0) / G, (self.y1 - self.y0) / G, (self.z1 - self.z0) / G
        xs = self.x0 + (np.arange(G) + 0.5) * hx
        ys = self.y0 + (np.arange(G) + 0.5) * hy
        zs = self.z0 + (np.arange(G) + 0.5) * hz
        return xs, ys, zs


def _cell_keys(model: Model):
    """Integer (level, ix, iy, iz) keys of every cell, sorted, with cell ids."""
    mp = model.mesh_params
    dom = np.asarray(mp.domain, np.float64)
    lev = model.array("cell_level").astype(np.int64)
    cx, cy, cz = model.array("cell_cx"), model.array("cell_cy"), model.array("cell_cz")
    n = (1 << lev).astype(np.float64)
    ix = np.floor((cx - dom[0]) / (dom[1] - dom[0]) * n).astype(np.int64)
    iy = np.floor((cy - dom[2]) / (dom[3] - dom[2]) * n).astype(np.int64)
    iz = np.floor((cz - dom[4]) / (dom[5] - dom[4]) * n).astype(np.int64)
    key = (lev << 57) | (ix << 38) | (iy << 19) | iz
    order = np.argsort(key)
    return key[order], order, np.unique(lev)


def sample_sigma(model: Model, sigma: np.ndarray, box: Box) -> np.ndarray:
    """log10(sigma) at the voxel centres, from the cell containing each centre
    ([G, G, G] in z, y, x order)."""
    mp = model.mesh_params
    dom = np.asarray(mp.domain, np.float64)
    keys, order, levels = _cell_keys(model)
    xs, ys, zs = box.centres()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    px, py, pz = X.ravel(), Y.ravel(), Z.ravel()
    out = np.full(px.shape, np.log10(configs.SIGMA_AIR))
    found = np.zeros(px.shape, bool)
    for L in levels:
        n = float(1 << int(L))
        ix = np.floor((px - dom[0]) / (dom[1] - dom[0]) * n).astype(np.int64)
        iy = np.floor((py - dom[2]) / (dom[3] - dom[2]) * n).astype(np.int64)
        iz = np.floor((pz - dom[4]) / (dom[5] - dom[4]) * n).astype(np.int64)
        q = (np.int64(L) << 57) | (ix << 38) | (iy << 19) | iz
        pos = np.clip(np.searchsorted(keys, q), 0, len(keys) - 1)
        hit = (keys[pos] == q) & ~found
        out[hit] = np.log10(sigma[order[pos[hit]]])
        found |= hit
    return out.reshape(G, G, G)


def features(logsig: np.ndarray, freqs) -> np.ndarray:
    """Network input [nf, G^3, 2]: scaled log10(sigma) and log10(f)."""
    freqs = np.asarray(freqs, np.float64)
    x = np.zeros((len(freqs), G * G * G, 2), np.float32)
    x[:, :, 0] = ((logsig.ravel() + 2.0) / 3.0)[None, :]
    x[:, :, 1] = (np.log10(freqs) / 2.5)[:, None]
    return x


def deposit_weights(model: Model, box: Box):
    """Trilinear deposition of a model's edges onto the box's voxel centres:
    per corner c of the 8, the voxel index and weight of every edge inside
    the box, with the edges' direction and indices (the geometry part of
    deposit, shared by every field of the model)."""
    ex, ey, ez = model.array("edge_x"), model.array("edge_y"), model.array("edge_z")
    d = model.array("edge_dir").astype(np.int64)
    hx, hy, hz = (box.x1 - box.x0) / G, (box.y1 - box.y0) / G, (box.z1 - box.z0) / G
    fx, fy, fz = (ex - box.x0) / hx - 0.5, (ey - box.y0) / hy - 0.5, (ez - box.z0) / hz - 0.5
    inside = (fx >= 0) & (fy >= 0) & (fz >= 0) & (fx <= G - 1) & (fy <= G - 1) & (fz <= G - 1)
    sel = np.flatnonzero(inside)
    fx, fy, fz, d = fx[sel], fy[sel], fz[sel], d[sel]
    ix = np.minimum(fx.astype(np.int64), G - 2)
    iy = np.minimum(fy.astype(np.int64), G - 2)
    iz = np.minimum(fz.astype(np.int64), G - 2)
    tx, ty, tz = fx - ix, fy - iy, fz - iz
    vox, wts = [], []
    for c in range(8):
        cx, cy, cz = c & 1, (c >> 1) & 1, c >> 2
        wts.append((tx if cx else 1 - tx) * (ty if cy else 1 - ty) * (tz if cz else 1 - tz))
        vox.append(((iz + cz) * G + iy + cy) * G + ix + cx)
    return sel, d, np.stack(vox), np.stack(wts)


def deposit(model: Model, box: Box, e_sec_pol1: np.ndarray, e_sec_pol2: np.ndarray,
            geometry=None) -> np.ndarray:
    """Trilinear deposition of edge values onto the voxel centres ([G^3, 12]:
    channel pol*6 + 2*dir + {re, im}): per voxel and direction, the
    weight-averaged values of the edges of that direction around it (0 where
    there is none).  geometry: deposit_weights of the model, when reused."""
    sel, d, vox, wts = geometry if geometry is not None else deposit_weights(model, box)
    n3 = G * G * G
    out = np.zeros((n3, NCH))
    vals = (e_sec_pol1[sel], e_sec_pol2[sel])
    for dd in range(3):
        m = d == dd
        if not m.any():
            continue
        v_ = vox[:, m].ravel()
        w_ = wts[:, m].ravel()
        wsum = np.bincount(v_, weights=w_, minlength=n3)
        ok = wsum > 0
        for pol in range(2):
            v = np.tile(vals[pol][m], 8)
            for k, part in enumerate((v.real, v.imag)):
                acc = np.bincount(v_, weights=w_ * part, minlength=n3)
                out[ok, pol * 6 + 2 * dd + k] = acc[ok] / wsum[ok]
    return out.astype(np.float32)


class CoarseNet:
    """Device network (include/m3d_net.h)."""

    def __init__(self, ctx: Context, max_batch: int = 8, dilations=(1, 2, 4, 1), seed: int = 0):
        L = lib()
        vp, i, f = C.c_void_p, C.c_int, C.c_float
        L.m3d_net_last_error.restype = C.c_char_p
        L.m3d_net_create.argtypes = [vp, i, vp, C.c_uint, C.POINTER(vp)]
        L.m3d_net_destroy.argtypes = [vp]
        L.m3d_net_num_params.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.m3d_net_get_params.argtypes = [vp, vp, vp, vp, vp]
        L.m3d_net_set_params.argtypes = [vp, vp, vp]
        L.m3d_net_forward.argtypes = [vp, i, vp, vp]
        L.m3d_net_gradients.argtypes = [vp, i, vp, vp, C.POINTER(f)]
        L.m3d_net_train_step.argtypes = [vp, i, vp, vp, f, C.POINTER(f)]
        L.m3d_net_set_clip.argtypes = [vp, f]
        L.m3d_net_prolong.argtypes = [vp, i, vp, i, vp, vp, vp, vp]
        self.L = L
        self.ctx = ctx
        self.max_batch = max_batch
        self.dil = tuple(dilations)
        d = (C.c_int * 4)(*dilations)
        h = C.c_void_p()
        self._chk(L.m3d_net_create(ctx.handle, max_batch, d, seed, C.byref(h)))
        self.handle = h
        nw, nb = C.c_longlong(), C.c_longlong()
        L.m3d_net_num_params(C.byref(nw), C.byref(nb))
        self.nw, self.nb = nw.value, nb.value

    def _chk(self, st):
        if st != 0:
            raise NativeError(self.L.m3d_net_last_error().decode())

    def close(self):
        if getattr(self, "handle", None):
            self.L.m3d_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, grads: bool = False):
        w = np.zeros(self.nw, np.float32)
        b = np.zeros(self.nb, np.float32)
        if grads:
            self._chk(self.L.m3d_net_get_params(self.handle, None, None, _ptr(w), _ptr(b)))
        else:
            self._chk(self.L.m3d_net_get_params(self.handle, _ptr(w), _ptr(b), None, None))
        return w, b

    def set_params(self, w, b):
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        self._chk(self.L.m3d_net_set_params(self.handle, _ptr(w), _ptr(b)))

    def forward(self, d_in, d_out, batch: int):
        self._chk(self.L.m3d_net_forward(self.handle, batch, C.c_void_p(d_in), C.c_void_p(d_out)))

    def gradients(self, d_in, d_tgt, batch: int) -> float:
        loss = C.c_float()
        self._chk(self.L.m3d_net_gradients(self.handle, batch, C.c_void_p(d_in), C.c_void_p(d_tgt),
                                           C.byref(loss)))
        return loss.value

    def set_clip(self, max_norm: float) -> None:
        """Global gradient-norm clip applied by train_step (0 disables)."""
        self._chk(self.L.m3d_net_set_clip(self.handle, max_norm))

    def train_step(self, d_in, d_tgt, batch: int, lr: float) -> float:
        loss = C.c_float()
        self._chk(self.L.m3d_net_train_step(self.handle, batch, C.c_void_p(d_in), C.c_void_p(d_tgt),
                                            lr, C.byref(loss)))
        return loss.value

    def prolong(self, d_field, batch: int, n_edges: int, d_exyz, d_edir, box: Box, d_x0):
        bx = box.arr()
        self._chk(self.L.m3d_net_prolong(self.handle, batch, C.c_void_p(d_field), n_edges,
                                         C.c_void_p(d_exyz), C.c_void_p(d_edir), _ptr(bx),
                                         C.c_void_p(d_x0)))


# --------------------------------------------------------------------------
# Data generation, training and evaluation (torch tensors are device memory
# plumbing only).
# --------------------------------------------------------------------------
@dataclass
class Sample:
    x: np.ndarray       # [nf, G^3, 2]
    y: np.ndarray       # [nf, G^3, 12]
    iterations: list = field(default_factory=list)


def solve_model(model: Model, wl: configs.Workload, params=None):
    """Forward solve of every frequency; returns E_sec per system and results.
    Training targets need far less than the solve tolerance, so by default the
    mixed-precision solve with interleaved DivCorr runs to a relative residual
    of 1e-6."""
    if params is None:
        params = wl.fwd.fast_params(tol=1e-6)
    E, Z, res = model.solve_frequencies(wl.freqs, params)
    esec = np.empty_like(E)
    for f, fr in enumerate(wl.freqs):
        w = 2 * np.pi * fr
        for pol in range(2):
            esec[2 * f + pol] = E[2 * f + pol] - model.primary_field(w, pol)
    return esec, res


def normalise(y: np.ndarray) -> np.ndarray:
    """Scale each polarisation's six channels of each sample ([nf, G^3, 12])
    to unit RMS.  The network learns the shape of E_sec; the forward solve
    restores the amplitude per system with its least-squares scaling of the
    guess (use_initial_guess = 2)."""
    out = y.copy()
    for pol in range(2):
        ch = slice(6 * pol, 6 * pol + 6)
        rms = np.sqrt(np.mean(out[:, :, ch].astype(np.float64) ** 2, axis=(1, 2)))
        rms[rms == 0] = 1.0
        out[:, :, ch] /= rms[:, None, None].astype(np.float32)
    return out


def make_sample(model: Model, wl: configs.Workload, box: Box) -> Optional[Sample]:
    """Training sample of one model (None if a solve did not converge)."""
    esec, res = solve_model(model, wl)
    if not all(r["converged"] for r in res):
        return None
    return sample_from_fields(model, wl, box, esec, res)


def sample_from_fields(model: Model, wl: configs.Workload, box: Box, esec: np.ndarray, res) -> Sample:
    """Training sample from a model's scattered fields [2 nf, n]."""
    logsig = sample_sigma(model, model.array("sigma"), box)
    x = features(logsig, wl.freqs)
    geo = deposit_weights(model, box)
    y = np.stack([deposit(model, box, esec[2 * f], esec[2 * f + 1], geo) for f in range(len(wl.freqs))])
    return Sample(x, normalise(y), [r["iterations"] for r in res])


def training_workload(seed: int, max_level: int = 5) -> configs.Workload:
    return configs.random3d(seed, max_level=max_level, half_width=100e3, station_radius=20e3)


def seeded_level(seed: int, levels=(5, 6, 7), weights=(0.5, 0.35, 0.15)) -> int:
    """Finest octree level of the model of one seed, drawn from `levels` with
    `weights` (the same level for a seed every time)."""
    rng = np.random.default_rng(seed + 7_777_777)
    w = np.asarray(weights, np.float64)
    return int(levels[int(rng.choice(len(levels), p=w / w.sum()))])


def domain_box(wl: configs.Workload) -> Box:
    """The voxel box over the whole mesh domain."""
    return Box(*[float(v) for v in wl.mesh.domain])


def train(net: CoarseNet, samples: List[Sample], steps: int, lr: float = 5e-4, batch: int = 8,
          seed: int = 0, log=None, warmup: int = 100) -> List[float]:
    """Adam on random mini-batches: linear warm-up over `warmup` steps, then
    cosine decay to lr/100; gradients clipped to twice the median global norm
    of the first 20 batches."""
    import torch
    dev = torch.device("cuda:0")
    X = torch.from_numpy(np.concatenate([s.x for s in samples])).to(dev)
    Y = torch.from_numpy(np.concatenate([s.y for s in samples])).to(dev)
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    xb = torch.empty((batch,) + tuple(X.shape[1:]), dtype=torch.float32, device=dev)
    yb = torch.empty((batch,) + tuple(Y.shape[1:]), dtype=torch.float32, device=dev)

    def load():
        idx = torch.from_numpy(rng.integers(0, n, batch)).to(dev)
        xb.copy_(X[idx])
        yb.copy_(Y[idx])
        torch.cuda.synchronize()

    norms = []
    for _ in range(20):
        load()
        net.gradients(xb.data_ptr(), yb.data_ptr(), batch)
        gw, gb = net.params(grads=True)
        norms.append(float(np.sqrt(np.sum(gw.astype(np.float64) ** 2) + np.sum(gb.astype(np.float64) ** 2))))
    clip = 2.0 * float(np.median(norms))
    net.set_clip(clip)
    if log:
        log(f"[learned] gradient clip {clip:.3e} (median initial norm {np.median(norms):.3e})")
    losses = []
    for it in range(steps):
        load()
        if it < warmup:
            step_lr = lr * (it + 1) / warmup
        else:
            t = (it - warmup) / max(steps - warmup, 1)
            step_lr = lr * (0.01 + 0.99 * 0.5 * (1.0 + np.cos(np.pi * t)))
        losses.append(net.train_step(xb.data_ptr(), yb.data_ptr(), batch, step_lr))
        if log and (it % 500 == 0 or it == steps - 1):
            log(f"[learned] step {it} lr {step_lr:.2e} loss {losses[-1]:.4e} "
                f"(mean of last 100: {np.mean(losses[-100:]):.4e})")
    return losses


def initial_guess(net: CoarseNet, model: Model, wl: configs.Workload, box: Box):
    """Device e0 for every (frequency, polarisation) system of wl on model:
    returns a torch complex128 tensor [2*nf, n] (system-major, edge order)."""
    import torch
    dev = torch.device("cuda:0")
    nf = len(wl.freqs)
    logsig = sample_sigma(model, model.array("sigma"), box)
    x = torch.from_numpy(features(logsig, wl.freqs)).to(dev)
    out = torch.empty((nf, G * G * G, NCH), dtype=torch.float32, device=dev)
    exyz = torch.from_numpy(np.stack([model.array("edge_x"), model.array("edge_y"),
                                      model.array("edge_z")], 1).astype(np.float32)).to(dev)
    edir = torch.from_numpy(model.array("edge_dir")).to(dev)
    x0 = torch.empty((2 * nf, model.n), dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    for s in range(0, nf, net.max_batch):
        b = min(net.max_batch, nf - s)
        net.forward(x[s].data_ptr(), out[s].data_ptr(), b)
        net.prolong(out[s].data_ptr(), b, model.n, exyz.data_ptr(), edir.data_ptr(), box,
                    x0[2 * s].data_ptr())
    return x0
