"""Gauss-Newton Hessian-vector product of the MT misfit on the device.

The data term of ImplicitJtJOperatorFV::Mult (src/inversion/implicit_jtj_fv.cpp:78-192):
for x = delta ln(sigma) per cell,

    y = sum over frequencies f and polarisations p of
        S_fp( Q_fp^T( W^2 dZ_fp ) ),   dZ_fp = D_p( A_f^-1 P_fp(sigma x) ),

where P_fp is the perturbation right-hand side with the reference's
harmonic-mean L^3 mass derivative (build_perturbation_rhs,
forward_solver_fv.cpp:945-984), D_p the station delta-Z of one polarisation
(extract_delta_impedance_complex, :2018-2065), Q_fp^T the adjoint right-hand
side of that polarisation (build_adjoint_rhs_from_residual, :1804-2013; the
weighted delta Z enters unconjugated, as in the reference), and S_fp the
sensitivity with that polarisation's background field (compute_sensitivity,
:989-1059).  Both solves run with the inner tolerance 1e-4 of
set_solver_tolerance_override and interleaved DivCorr (solve_rhs_complex /
adjoint_solve_complex, :914-940).  The regularisation term lambda W^T W x of
Mult belongs to the inversion driver, which is not part of this path.

Station H.  delta Z and the adjoint right-hand side divide by the station
H of the background field.  The reference's Mult never refreshes its station
cache between frequencies (only an impedance extraction does, and the
background pass ends on the last frequency, implicit_jtj_fv.cpp:37-55,
117-119), so every frequency uses the H of the last one; h_cache="reference"
(the default) does the same, h_cache="per_frequency" uses each frequency's
own H.  The reference's inner solves also run on the operator as
factorize_frequency leaves it, before the Dirichlet / dead-edge elimination;
here they run on the eliminated operator, the only one the device keeps.

Instead of the reference's loop over frequencies and polarisations, every
perturbation system runs in one device batch and every adjoint system in a
second.  The fields stay on the device (the perturbation right-hand sides
are formed there, m3d_perturbation_rhs_device); only the station cells'
edges cross to the host, for delta Z and the adjoint right-hand sides.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .native import MODE_FORWARD, MODE_INTERLEAVED, Model, SolveParams


class GaussNewton:
    """J^T W^2 J of one model at fixed sigma (the background solve runs once,
    in the constructor)."""

    def __init__(self, model: Model, freqs, z_err: np.ndarray, params: Optional[SolveParams] = None,
                 inner_tol: float = 1e-4, h_cache: str = "reference"):
        if h_cache not in ("reference", "per_frequency"):
            raise ValueError("h_cache is 'reference' or 'per_frequency'")
        self.h_cache = h_cache
        import torch
        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.model = model
        self.freqs = np.asarray(freqs, dtype=np.float64)
        nf = len(self.freqs)
        self.omegas = np.repeat(2 * np.pi * self.freqs, 2)
        err = np.asarray(z_err)
        self.w2 = np.where(err > 0, 1.0 / np.where(err > 0, err, 1.0) ** 2, 0.0)   # [nf, ns, 4]
        fwd = params or model.fwd_params.solve_params()
        self.inner = SolveParams(**{**fwd.__dict__, "tol": inner_tol})
        rhs = np.stack([model.rhs_total(model.bc_values(w, p % 2)) for p, w in enumerate(self.omegas)])
        aux = np.stack([model.primary_field(w, p % 2) for p, w in enumerate(self.omegas)])
        d_rhs = torch.from_numpy(rhs).to(self.dev)
        d_aux = torch.from_numpy(aux).to(self.dev)
        self.d_ebg = torch.zeros_like(d_rhs)
        torch.cuda.synchronize()
        self.background = model.device.solve_device(MODE_FORWARD, self.omegas, d_rhs.data_ptr(),
                                                    d_aux.data_ptr(), self.d_ebg.data_ptr(), 2 * nf, fwd)
        self.idx = torch.from_numpy(model.station_edges().astype(np.int64)).to(self.dev)
        self.ebg_c = self.d_ebg.index_select(1, self.idx).cpu().numpy()
        self.sigma = model.array("sigma").copy()

    def _perturbation_fields(self, x):
        """delta E of every (frequency, polarisation) system, on the device."""
        torch = self.torch
        d_ds = torch.from_numpy(self.sigma * np.asarray(x, dtype=np.float64)).to(self.dev)
        d_prhs = torch.empty_like(self.d_ebg)
        d_de = torch.zeros_like(d_prhs)
        torch.cuda.synchronize()
        self.model.device.perturbation_rhs_device(self.omegas, self.d_ebg.data_ptr(), d_ds.data_ptr(),
                                                  d_prhs.data_ptr())
        res = self.model.device.solve_device(MODE_INTERLEAVED, self.omegas, d_prhs.data_ptr(), 0,
                                             d_de.data_ptr(), len(self.omegas), self.inner)
        return d_de, res

    def _station_h(self, f):
        """Load the station H cache used at frequency f: that of the last
        background frequency (h_cache "reference") or of f itself."""
        g = len(self.freqs) - 1 if self.h_cache == "reference" else f
        self.model.extract_impedance_compact(self.omegas[2 * g], self.ebg_c[2 * g], self.ebg_c[2 * g + 1])

    def _delta_z(self, dec):
        """[nf, 2, stations, 4] delta Z per polarisation from compact delta E."""
        m = self.model
        nf = len(self.freqs)
        out = np.zeros((nf, 2, m.count("stations"), 4), np.complex128)
        for f in range(nf):
            self._station_h(f)
            for p in range(2):
                out[f, p] = m.delta_impedance_compact(dec[2 * f + p], p + 1)
        return out

    def jvp(self, x):
        """(delta Z per polarisation [nf, 2, stations, 4], delta E on the
        device [2 nf, n]) for x = delta ln(sigma) per cell."""
        d_de, _ = self._perturbation_fields(x)
        dec = d_de.index_select(1, self.idx).cpu().numpy()
        return self._delta_z(dec), d_de

    def matvec(self, x) -> np.ndarray:
        """y = J^T W^2 J x per cell (restrict to the active cells outside)."""
        torch = self.torch
        m = self.model
        nf = len(self.freqs)
        dz, _ = self.jvp(x)
        k = len(m.station_edges())
        arhs = np.zeros((2 * nf, k), np.complex128)
        for f in range(nf):
            w = self.omegas[2 * f]
            self._station_h(f)
            for p in range(2):
                r1, r2 = m.adjoint_rhs_compact(w, self.w2[f] * dz[f, p])
                arhs[2 * f + p] = r1 if p == 0 else r2
        d_arhs = torch.zeros_like(self.d_ebg)
        d_arhs.index_copy_(1, self.idx, torch.from_numpy(arhs).to(self.dev))
        d_lam = torch.zeros_like(d_arhs)
        torch.cuda.synchronize()
        m.device.solve_device(MODE_INTERLEAVED, self.omegas, d_arhs.data_ptr(), 0, d_lam.data_ptr(),
                              len(self.omegas), self.inner)
        d_y = torch.empty(m.nc, dtype=torch.float64, device=self.dev)
        m.device.sensitivity_device(self.omegas, self.d_ebg.data_ptr(), d_lam.data_ptr(), d_y.data_ptr())
        return d_y.cpu().numpy()


def matvec_host(model: Model, freqs, z_err, x, params: Optional[SolveParams] = None,
                inner_tol: float = 1e-4, h_cache: str = "reference") -> np.ndarray:
    """The same product in the reference's order — per frequency and
    polarisation, host buffers, one system at a time — for checking the
    batched operator."""
    freqs = np.asarray(freqs, dtype=np.float64)
    err = np.asarray(z_err)
    w2 = np.where(err > 0, 1.0 / np.where(err > 0, err, 1.0) ** 2, 0.0)
    fwd = params or model.fwd_params.solve_params()
    inner = SolveParams(**{**fwd.__dict__, "tol": inner_tol})
    E, _, _ = model.solve_frequencies(freqs, fwd)
    sigma = model.array("sigma")
    ds = sigma * np.asarray(x, dtype=np.float64)
    idx = model.station_edges()
    y = np.zeros(model.nc)
    last = len(freqs) - 1
    for f, fr in enumerate(freqs):
        w = 2 * np.pi * fr
        g = last if h_cache == "reference" else f
        model.extract_impedance(2 * np.pi * freqs[g], E[2 * g], E[2 * g + 1])
        for p in range(2):
            prhs = model.perturbation_rhs(w, E[2 * f + p], ds)
            de, _ = model.device.solve(MODE_INTERLEAVED, [w], prhs[None, :], params=inner)
            dz = model.delta_impedance_compact(de[0][idx], p + 1)
            r1, r2 = model.adjoint_rhs(w, w2[f] * dz)
            lam, _ = model.device.solve(MODE_INTERLEAVED, [w], (r1 if p == 0 else r2)[None, :], params=inner)
            y += model.sensitivity(w, E[2 * f + p], lam[0])
    return y
