"""Workload definitions: meshes, stations, frequencies and conductivity models.

* ``halfspace`` -- uniform 100 ohm-m halfspace under a 1e-8 S/m air layer,
  domain half-width 50 km (z in [-50, 50] km, cubic cells), octree levels 3/5,
  5x5 surface stations at 2 km spacing, 1 Hz, both polarisations.
* ``random3d`` -- seeded random 3-D model: 3 layers with log-uniform 10-1000
  ohm-m resistivity and 0.5-20 km thicknesses (the last a halfspace) plus 1-4
  axis-aligned blocks of log-uniform 1-10^4 ohm-m with 1-10 km sides in the
  upper 15 km; octree levels 3-7; 11x11 stations at 1 km spacing; 8
  frequencies log-spaced 1e-3..1e2 Hz.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .native import Context, ForwardParams, MeshParams, Model

SIGMA_AIR = 1e-8


@dataclass
class Workload:
    name: str
    mesh: MeshParams
    stations: np.ndarray          # [ns, 3]
    freqs: np.ndarray             # Hz
    fwd: ForwardParams
    layers: List[tuple] = field(default_factory=list)   # (z_top, z_bot, sigma)
    blocks: List[tuple] = field(default_factory=list)   # (x0,x1,y0,y1,z0,z1,sigma)
    sigma_bg: float = 0.01
    seed: Optional[int] = None


def station_grid(n: int, spacing: float) -> np.ndarray:
    c = (np.arange(n) - (n - 1) / 2.0) * spacing
    xx, yy = np.meshgrid(c, c, indexing="xy")
    return np.stack([xx.ravel(), yy.ravel(), np.zeros(n * n)], axis=1)


def halfspace(max_level: int = 5, divcorr_iter_per_dc: int = 100) -> Workload:
    st = station_grid(5, 2000.0)
    mesh = MeshParams(domain=(-50e3, 50e3, -50e3, 50e3, -50e3, 50e3), min_level=3,
                      max_level=max_level, sigma_bg=0.01)
    fwd = ForwardParams(bicgstab_tol=1e-7, divcorr_iter_per_dc=divcorr_iter_per_dc)
    return Workload("halfspace_L3-5_5x5st_1Hz", mesh, st, np.array([1.0]), fwd,
                    layers=[(0.0, -np.inf, 0.01)], sigma_bg=0.01)


def random3d(seed: int, min_level: int = 3, max_level: int = 7, n_freq: int = 8,
             half_width: float = 100e3, station_radius: float = 20e3,
             maxiter: int = 20000) -> Workload:
    rng = np.random.default_rng(seed)
    rho_layers = 10.0 ** rng.uniform(1.0, 3.0, size=3)
    thick = rng.uniform(500.0, 20e3, size=2)
    z1 = -thick[0]
    z2 = z1 - thick[1]
    layers = [(0.0, z1, 1.0 / rho_layers[0]), (z1, z2, 1.0 / rho_layers[1]),
              (z2, -np.inf, 1.0 / rho_layers[2])]
    nb = int(rng.integers(1, 5))
    blocks = []
    for _ in range(nb):
        sides = rng.uniform(1e3, 10e3, size=3)
        cx, cy = rng.uniform(-10e3, 10e3, size=2)
        ztop = -rng.uniform(0.0, 15e3 - sides[2])
        rho = 10.0 ** rng.uniform(0.0, 4.0)
        blocks.append((cx - sides[0] / 2, cx + sides[0] / 2, cy - sides[1] / 2, cy + sides[1] / 2,
                       ztop - sides[2], ztop, 1.0 / rho))
    st = station_grid(11, 1000.0)
    # Background for the scattered-field primary: the top layer.
    sigma_bg = layers[0][2]
    mesh = MeshParams(domain=(-half_width, half_width, -half_width, half_width, -half_width,
                              half_width), min_level=min_level, max_level=max_level,
                      station_refine_radius=station_radius, station_refine_level=max_level,
                      sigma_bg=sigma_bg,
                      station_bbox=(st[:, 0].min(), st[:, 0].max(), st[:, 1].min(), st[:, 1].max()))
    freqs = np.logspace(-3.0, 2.0, n_freq)
    fwd = ForwardParams(bicgstab_tol=1e-7, bicgstab_maxiter=maxiter)
    return Workload(f"random3d_seed{seed}_L{min_level}-{max_level}", mesh, st, freqs, fwd,
                    layers=layers, blocks=blocks, sigma_bg=sigma_bg, seed=seed)


def conductivity(model: Model, wl: Workload) -> np.ndarray:
    """Cell conductivity: air 1e-8 S/m above the surface, layers by cell-centre
    depth, blocks by cell-centre containment."""
    cz = model.array("cell_cz")
    cx = model.array("cell_cx")
    cy = model.array("cell_cy")
    ctype = model.array("cell_type")
    sigma = np.full(model.nc, SIGMA_AIR)
    earth = ctype != 1
    for (zt, zb, s) in wl.layers:
        sel = earth & (cz <= zt) & (cz > zb)
        sigma[sel] = s
    for (x0, x1, y0, y1, z0, z1, s) in wl.blocks:
        sel = earth & (cx >= x0) & (cx <= x1) & (cy >= y0) & (cy <= y1) & (cz >= z0) & (cz <= z1)
        sigma[sel] = s
    return sigma


def build_model(wl: Workload, ctx: Optional[Context] = None) -> Model:
    m = Model(wl.mesh, wl.stations, wl.freqs, wl.fwd, ctx)
    m.set_sigma(conductivity(m, wl))
    m.find_stations(wl.stations[:, :2])
    return m
