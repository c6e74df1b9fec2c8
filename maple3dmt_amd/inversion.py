"""Data misfit and its adjoint gradient in ln(sigma), batched on the device.

Follows InversionFV::compute_gradient_ (src/inversion/inversion_fv.cpp:230-312):
per frequency a forward solve of both polarisations, the weighted residual
w^2 conj(d_obs - d_pred) (:258-271), the adjoint right-hand side
(build_adjoint_rhs_from_residual, forward_solver_fv.cpp:1804-2013), two
adjoint solves with interleaved DivCorr and the sensitivity of both
polarisations (:989-1059).  Instead of looping over frequencies, the device
runs two phases: every frequency's forward systems in one batch, then every
adjoint system in one batch; the forward -> adjoint dependency is the only
synchronisation.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .native import MODE_FORWARD, MODE_INTERLEAVED, Model, SolveParams


@dataclass
class GradientResult:
    phi: float                 # 1/2 sum w^2 |d_obs - d_pred|^2
    grad: np.ndarray           # d phi / d ln(sigma_c), per cell
    z_pred: np.ndarray         # [nf, stations, 4]
    forward: list
    adjoint: list


def weighted_residual(obs: np.ndarray, pred: np.ndarray, err: np.ndarray,
                      weight: Optional[np.ndarray] = None) -> np.ndarray:
    """w^2 conj(d_obs - d_pred) with w = 1/error; zero where error or weight
    is not positive (inversion_fv.cpp:258-271)."""
    wgt = np.ones_like(err) if weight is None else weight
    ok = (err > 0) & (wgt > 0)
    out = np.zeros_like(obs)
    out[ok] = np.conj(obs[ok] - pred[ok]) / err[ok] ** 2
    return out


def misfit(obs, pred, err, weight=None) -> float:
    wgt = np.ones_like(err) if weight is None else weight
    ok = (err > 0) & (wgt > 0)
    return 0.5 * float(np.sum(np.abs(obs[ok] - pred[ok]) ** 2 / err[ok] ** 2))


def gradient(model: Model, freqs, z_obs: np.ndarray, z_err: np.ndarray,
             params: Optional[SolveParams] = None,
             adjoint_params: Optional[SolveParams] = None) -> GradientResult:
    freqs = np.asarray(freqs, dtype=np.float64)
    nf = len(freqs)
    E, Z, fres = model.solve_frequencies(freqs, params)
    phi = misfit(z_obs, Z, z_err)
    rhs = []
    omegas = []
    for f in range(nf):
        w = 2 * np.pi * freqs[f]
        model.extract_impedance(w, E[2 * f], E[2 * f + 1])  # station H cache for this frequency
        wr = weighted_residual(z_obs[f], Z[f], z_err[f])
        r1, r2 = model.adjoint_rhs(w, wr)
        rhs += [r1, r2]
        omegas += [w, w]
    lam, ares = model.adjoint_solve(np.array(omegas), np.stack(rhs), adjoint_params or params)
    g = np.zeros(model.nc)
    for f in range(nf):
        w = omegas[2 * f]
        g += model.sensitivity(w, E[2 * f], lam[2 * f])
        g += model.sensitivity(w, E[2 * f + 1], lam[2 * f + 1])
    return GradientResult(phi, g, Z, fres, ares)


def gradient_device(model: Model, freqs, z_obs: np.ndarray, z_err: np.ndarray,
                    params: Optional[SolveParams] = None,
                    adjoint_params: Optional[SolveParams] = None) -> GradientResult:
    """gradient() with the fields resident on the device: the forward batch
    (FORWARD), the adjoint batch (INTERLEAVED) and the sensitivity sum over
    every frequency and polarisation (m3d_sensitivity_device) run on the GPU.
    Only the station cells' edges cross to the host: E is gathered there on
    the device, Z and the station right-hand sides come from the compact
    values (the reference's formulas, equal to the full-field path), and the
    right-hand sides are scattered into zeroed device vectors.  torch
    tensors are the device buffers."""
    import torch
    dev = torch.device("cuda:0")
    freqs = np.asarray(freqs, dtype=np.float64)
    nf = len(freqs)
    omegas = np.repeat(2 * np.pi * freqs, 2)
    fwd_prm = params or model.fwd_params.solve_params()
    rhs = np.stack([model.rhs_total(model.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
    aux = np.stack([model.primary_field(w, p % 2) for p, w in enumerate(omegas)])
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_aux = torch.from_numpy(aux).to(dev)
    d_e = torch.zeros_like(d_rhs)
    torch.cuda.synchronize()
    fres = model.device.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(),
                                     d_e.data_ptr(), 2 * nf, fwd_prm)
    idx = torch.from_numpy(model.station_edges().astype(np.int64)).to(dev)
    Ec = d_e.index_select(1, idx).cpu().numpy()          # [2 nf, station edges]
    Z = np.zeros_like(z_obs)
    arhs_c = np.zeros_like(Ec)
    for f in range(nf):
        w = omegas[2 * f]
        Z[f] = model.extract_impedance_compact(w, Ec[2 * f], Ec[2 * f + 1])
        arhs_c[2 * f], arhs_c[2 * f + 1] = model.adjoint_rhs_compact(
            w, weighted_residual(z_obs[f], Z[f], z_err[f]))
    phi = misfit(z_obs, Z, z_err)
    d_arhs = torch.zeros_like(d_e)
    d_arhs.index_copy_(1, idx, torch.from_numpy(arhs_c).to(dev))
    d_lam = torch.zeros_like(d_arhs)
    torch.cuda.synchronize()
    ares = model.device.solve_device(MODE_INTERLEAVED, omegas, d_arhs.data_ptr(), 0, d_lam.data_ptr(),
                                     2 * nf, adjoint_params or fwd_prm)
    d_g = torch.empty(model.nc, dtype=torch.float64, device=dev)
    model.device.sensitivity_device(omegas, d_e.data_ptr(), d_lam.data_ptr(), d_g.data_ptr())
    return GradientResult(phi, d_g.cpu().numpy(), Z, fres, ares)
