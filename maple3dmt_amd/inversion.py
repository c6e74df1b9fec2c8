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

    @property
    def unconverged_forward(self) -> int:
        """Forward systems that ended without meeting the tolerance (stopped
        by a breakdown guard, the iteration cap or the stagnation rule)."""
        return sum(not r["converged"] for r in self.forward)

    @property
    def unconverged_adjoint(self) -> int:
        """Adjoint systems that ended without meeting the tolerance; the
        interleaved solve accepts a stagnated system as the reference does
        (forward_solver_fv.cpp:866-873), and its gradient is built on it."""
        return sum(not r["converged"] for r in self.adjoint)


def weighted_residual(obs: np.ndarray, pred: np.ndarray, err: np.ndarray,
                      weight: Optional[np.ndarray] = None) -> np.ndarray:
    """w w conj(d_obs - d_pred) with w = 1 / error; zero where the error or
    the weight is not positive (inversion_fv.cpp:258-271, the same
    operations)."""
    wgt = np.ones_like(err) if weight is None else weight
    ok = (err > 0) & (wgt > 0)
    out = np.zeros_like(obs)
    w = 1.0 / err[ok]
    out[ok] = (w * w) * np.conj(obs[ok] - pred[ok])
    return out


def misfit(obs, pred, err, weight=None) -> float:
    """sum of 0.5 w w |d_obs - d_pred|^2 over the used data
    (data_misfit_from_predicted_, inversion_fv.cpp:75-98)."""
    wgt = np.ones_like(err) if weight is None else weight
    ok = (err > 0) & (wgt > 0)
    r = obs[ok] - pred[ok]
    w = 1.0 / err[ok]
    return float(np.sum(0.5 * w * w * (r.real * r.real + r.imag * r.imag)))


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
                    adjoint_params: Optional[SolveParams] = None,
                    z_weight: Optional[np.ndarray] = None) -> GradientResult:
    """gradient() with everything between the two batches on the device: the
    forward batch (FORWARD), the impedance of every frequency
    (m3d_impedance_device), the weighted residual and misfit
    (m3d_weighted_residual_device: w w conj(d_obs - d_pred), inversion_fv.cpp:
    258-271, and the misfit in the order of :75-98), the adjoint right-hand
    sides (m3d_adjoint_rhs_device), the adjoint batch (INTERLEAVED) and the
    sensitivity sum over every frequency and polarisation
    (m3d_sensitivity_device).  No field or station value crosses the bus
    between the forward and the adjoint batch; Z, phi and the gradient are
    read back once at the end.  torch tensors are the device buffers."""
    import torch
    dev = torch.device("cuda:0")
    freqs = np.asarray(freqs, dtype=np.float64)
    nf = len(freqs)
    omegas = np.repeat(2 * np.pi * freqs, 2)
    fwd_prm = params or model.fwd_params.solve_params()
    rhs = np.stack([model.rhs_total(model.bc_values(w, p % 2)) for p, w in enumerate(omegas)])
    aux = np.stack([model.primary_field(w, p % 2) for p, w in enumerate(omegas)])
    st = model.device_stations(nf)
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_aux = torch.from_numpy(aux).to(dev)
    d_obs = torch.from_numpy(np.ascontiguousarray(z_obs, np.complex128)).to(dev)
    d_err = torch.from_numpy(np.ascontiguousarray(z_err, np.float64)).to(dev)
    d_wgt = (None if z_weight is None
             else torch.from_numpy(np.ascontiguousarray(z_weight, np.float64)).to(dev))
    d_e = torch.zeros_like(d_rhs)
    d_z = torch.zeros((nf, st.ns, 4), dtype=torch.complex128, device=dev)
    d_wr = torch.zeros_like(d_z)
    d_phi = torch.zeros(1, dtype=torch.float64, device=dev)
    d_arhs = torch.empty_like(d_e)
    d_lam = torch.zeros_like(d_e)
    d_g = torch.empty(model.nc, dtype=torch.float64, device=dev)
    # torch's stream only: a device-wide synchronize would also wait on (and
    # break the graph capture of) the streams other host threads are
    # solving on (multimodel.run).
    torch.cuda.current_stream().synchronize()
    fres = model.device.solve_device(MODE_FORWARD, omegas, d_rhs.data_ptr(), d_aux.data_ptr(),
                                     d_e.data_ptr(), 2 * nf, fwd_prm)
    st.impedance(2 * np.pi * freqs, d_e.data_ptr(), d_z.data_ptr())
    st.weighted_residual(nf, d_obs.data_ptr(), d_err.data_ptr(),
                         None if d_wgt is None else d_wgt.data_ptr(), d_wr.data_ptr(), d_phi.data_ptr())
    st.adjoint_rhs(nf, d_wr.data_ptr(), d_arhs.data_ptr())
    ares = model.device.solve_device(MODE_INTERLEAVED, omegas, d_arhs.data_ptr(), 0, d_lam.data_ptr(),
                                     2 * nf, adjoint_params or fwd_prm)
    model.device.sensitivity_device(omegas, d_e.data_ptr(), d_lam.data_ptr(), d_g.data_ptr())
    return GradientResult(float(d_phi.cpu().item()), d_g.cpu().numpy(), d_z.cpu().numpy(), fres, ares)
