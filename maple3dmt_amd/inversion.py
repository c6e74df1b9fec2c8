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

from .native import Model, SolveParams


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
