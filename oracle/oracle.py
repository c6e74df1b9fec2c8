"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use it as the checker.  The product (maple3dmt_amd) never
imports it.

Parity status of each oracle function (how it is pinned) is listed in the
module docstring of tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


class OrcGrid(C.Structure):
    _fields_ = [
        ("ne", C.c_int), ("nf", C.c_int), ("nc", C.c_int),
        ("edge_x", C.c_void_p), ("edge_y", C.c_void_p), ("edge_z", C.c_void_p),
        ("edge_len", C.c_void_p), ("edge_dir", C.c_void_p),
        ("edge_adj_ptr", C.c_void_p), ("edge_adj", C.c_void_p),
        ("face_x", C.c_void_p), ("face_y", C.c_void_p), ("face_z", C.c_void_p),
        ("face_area", C.c_void_p), ("face_normal", C.c_void_p),
        ("face_minus", C.c_void_p), ("face_plus", C.c_void_p),
        ("face_edges_ptr", C.c_void_p), ("face_edges", C.c_void_p),
        ("n_hanging_faces", C.c_int), ("hanging_faces", C.c_void_p),
        ("n_hanging_edges", C.c_int), ("hanging_edges", C.c_void_p),
        ("cell_edges_ptr", C.c_void_p), ("cell_edges", C.c_void_p),
        ("cell_cx", C.c_void_p), ("cell_cy", C.c_void_p), ("cell_cz", C.c_void_p),
        ("cell_level", C.c_void_p), ("cell_type", C.c_void_p),
        ("domain", C.c_double * 6), ("cell0_size", C.c_double), ("sigma_bg", C.c_double),
    ]


class OrcParams(C.Structure):
    _fields_ = [
        ("bicgstab_tol", C.c_double), ("bicgstab_maxiter", C.c_int),
        ("air_z_threshold", C.c_double), ("div_correction", C.c_int),
        ("div_corr_iters", C.c_int), ("div_corr_tol", C.c_double),
        ("divcorr_outer_max", C.c_int), ("divcorr_iter_per_dc", C.c_int),
    ]


class OrcResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("converged", C.c_int),
                ("info", C.c_int), ("outer", C.c_int), ("true_residual", C.c_double)]

    def as_dict(self):
        return dict(iterations=self.iterations, residual=self.residual,
                    converged=bool(self.converged), info=self.info, outer=self.outer,
                    true_residual=self.true_residual)


def build():
    """Compile liboracle.so from m3d_oracle.cpp (make)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcParams)]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_num_nodes.argtypes = [vp]
        L.orc_set_sigma.argtypes = [vp, vp]
        L.orc_assemble.argtypes = [vp, C.c_double]
        L.orc_get_csr.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_int)]
        L.orc_get_vec.argtypes = [vp, C.c_char_p, C.POINTER(vp), C.POINTER(C.c_int)]
        L.orc_spmv.argtypes = [vp, vp, vp]
        L.orc_ssor.argtypes = [vp, vp, vp]
        L.orc_bicgstab_csr.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(OrcResult)]
        L.orc_ssor_csr.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        L.orc_divcorr.argtypes = [vp, vp, vp]
        L.orc_solve_pol.argtypes = [vp, C.c_int, vp, C.POINTER(OrcResult)]
        L.orc_rhs.argtypes = [vp, C.c_int, vp, vp]
        L.orc_adjoint.argtypes = [vp, vp, vp, C.POINTER(OrcResult)]
        L.orc_find_stations.argtypes = [vp, vp, C.c_int, vp]
        L.orc_impedance.argtypes = [vp, vp, vp, vp]
        L.orc_adjoint_rhs.argtypes = [vp, vp, vp, vp]
        L.orc_sensitivity.argtypes = [vp, vp, vp, vp]
        L.orc_perturbation_rhs.argtypes = [vp, vp, vp, vp]
        L.orc_delta_impedance.argtypes = [vp, vp, C.c_int, vp]
        _lib = L
    return _lib


def _chk(st):
    if st != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Reference-shaped CPU solver built from a product Model's staggered grid."""

    def __init__(self, model, fwd=None, domain=None, sigma_bg=None):
        fwd = fwd or model.fwd_params
        mp = model.mesh_params
        self.keep = {}

        def arr(name, dtype):
            a = np.ascontiguousarray(model.array(name), dtype=dtype)
            self.keep[name] = a
            return _p(a)

        g = OrcGrid()
        g.ne, g.nf, g.nc = model.n, model.count("faces"), model.nc
        g.edge_x, g.edge_y, g.edge_z = arr("edge_x", np.float64), arr("edge_y", np.float64), arr("edge_z", np.float64)
        g.edge_len, g.edge_dir = arr("edge_len", np.float64), arr("edge_dir", np.uint8)
        g.edge_adj_ptr, g.edge_adj = arr("edge_adj_ptr", np.int32), arr("edge_adj", np.int32)
        g.face_x, g.face_y, g.face_z = arr("face_x", np.float64), arr("face_y", np.float64), arr("face_z", np.float64)
        g.face_area, g.face_normal = arr("face_area", np.float64), arr("face_normal", np.uint8)
        g.face_minus, g.face_plus = arr("face_minus", np.int32), arr("face_plus", np.int32)
        g.face_edges_ptr, g.face_edges = arr("face_edges_ptr", np.int32), arr("face_edges", np.int32)
        hf = np.ascontiguousarray(model.array("hanging_faces"), np.int32)
        he = np.ascontiguousarray(model.array("hanging_edges"), np.int32)
        self.keep["hf"], self.keep["he"] = hf, he
        g.n_hanging_faces, g.hanging_faces = len(hf) // 5, _p(hf)
        g.n_hanging_edges, g.hanging_edges = len(he) // 3, _p(he)
        g.cell_edges_ptr, g.cell_edges = arr("cell_edges_ptr", np.int32), arr("cell_edges", np.int32)
        g.cell_cx, g.cell_cy, g.cell_cz = arr("cell_cx", np.float64), arr("cell_cy", np.float64), arr("cell_cz", np.float64)
        g.cell_level, g.cell_type = arr("cell_level", np.int32), arr("cell_type", np.uint8)
        g.domain[:] = list(domain or mp.domain)
        g.cell0_size = float(model.array("cell_size")[0])
        g.sigma_bg = float(sigma_bg if sigma_bg is not None else mp.sigma_bg)
        p = OrcParams(fwd.bicgstab_tol, fwd.bicgstab_maxiter, fwd.air_z_threshold,
                      int(fwd.div_correction), fwd.div_corr_iters, fwd.div_corr_tol,
                      fwd.divcorr_outer_max, fwd.divcorr_iter_per_dc)
        self.handle = lib().orc_create(C.byref(g), C.byref(p))
        if not self.handle:
            raise RuntimeError(lib().orc_last_error().decode())
        self.ne, self.nc = model.n, model.nc
        self.nn = lib().orc_num_nodes(self.handle)
        self.ns = 0

    def __del__(self):
        try:
            if self.handle:
                lib().orc_destroy(self.handle)
        except Exception:
            pass

    def set_sigma(self, sigma):
        s = np.ascontiguousarray(sigma, np.float64)
        _chk(lib().orc_set_sigma(self.handle, _p(s)))

    def assemble(self, omega):
        _chk(lib().orc_assemble(self.handle, omega))

    def csr(self, name):
        nr, nz, cx = C.c_int(), C.c_int(), C.c_int()
        pp, cp, vp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _chk(lib().orc_get_csr(self.handle, name.encode(), C.byref(nr), C.byref(nz), C.byref(pp),
                               C.byref(cp), C.byref(vp), C.byref(cx)))
        ptr = np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_int)), (nr.value + 1,)).copy()
        col = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_int)), (max(nz.value, 1),))[:nz.value].copy()
        k = 2 if cx.value else 1
        val = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_double)), (max(k * nz.value, 1),))[:k * nz.value].copy()
        if cx.value:
            val = val.view(np.complex128)
        return ptr, col, val

    def vec(self, name):
        dp, n = C.c_void_p(), C.c_int()
        _chk(lib().orc_get_vec(self.handle, name.encode(), C.byref(dp), C.byref(n)))
        return np.ctypeslib.as_array(C.cast(dp, C.POINTER(C.c_double)), (max(n.value, 1),))[:n.value].copy()

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.ne, np.complex128)
        _chk(lib().orc_spmv(self.handle, _p(x), _p(y)))
        return y

    def ssor(self, r):
        r = np.ascontiguousarray(r, np.complex128)
        z = np.zeros(self.ne, np.complex128)
        _chk(lib().orc_ssor(self.handle, _p(r), _p(z)))
        return z

    def divcorr(self, E, phi0=None):
        E = np.ascontiguousarray(E, np.complex128).copy()
        p0 = None if phi0 is None else np.ascontiguousarray(phi0, np.complex128)
        _chk(lib().orc_divcorr(self.handle, _p(E), None if p0 is None else _p(p0)))
        return E

    def solve_pol(self, direction):
        E = np.zeros(self.ne, np.complex128)
        r = OrcResult()
        _chk(lib().orc_solve_pol(self.handle, direction, _p(E), C.byref(r)))
        return E, r.as_dict()

    def rhs(self, direction):
        rt = np.zeros(self.ne, np.complex128)
        ep = np.zeros(self.ne, np.complex128)
        _chk(lib().orc_rhs(self.handle, direction, _p(rt), _p(ep)))
        return rt, ep

    def adjoint(self, rhs):
        rhs = np.ascontiguousarray(rhs, np.complex128)
        x = np.zeros(self.ne, np.complex128)
        r = OrcResult()
        _chk(lib().orc_adjoint(self.handle, _p(rhs), _p(x), C.byref(r)))
        return x, r.as_dict()

    def find_stations(self, xy):
        a = np.ascontiguousarray(np.asarray(xy, np.float64).reshape(-1, 2))
        cells = np.zeros(a.shape[0], np.int32)
        _chk(lib().orc_find_stations(self.handle, _p(a), a.shape[0], _p(cells)))
        self.ns = a.shape[0]
        return cells

    def impedance(self, e1, e2):
        a = np.ascontiguousarray(e1, np.complex128)
        b = np.ascontiguousarray(e2, np.complex128)
        z = np.zeros((self.ns, 4), np.complex128)
        _chk(lib().orc_impedance(self.handle, _p(a), _p(b), _p(z)))
        return z

    def adjoint_rhs(self, wr):
        w = np.ascontiguousarray(wr, np.complex128)
        r1 = np.zeros(self.ne, np.complex128)
        r2 = np.zeros(self.ne, np.complex128)
        _chk(lib().orc_adjoint_rhs(self.handle, _p(w), _p(r1), _p(r2)))
        return r1, r2

    def sensitivity(self, e, lam):
        a = np.ascontiguousarray(e, np.complex128)
        b = np.ascontiguousarray(lam, np.complex128)
        g = np.zeros(self.nc, np.float64)
        _chk(lib().orc_sensitivity(self.handle, _p(a), _p(b), _p(g)))
        return g

    def perturbation_rhs(self, e_bg, dsigma):
        """build_perturbation_rhs at the assembled frequency (dsigma in sigma space)."""
        a = np.ascontiguousarray(e_bg, np.complex128)
        d = np.ascontiguousarray(dsigma, np.float64)
        r = np.zeros(self.ne, np.complex128)
        _chk(lib().orc_perturbation_rhs(self.handle, _p(a), _p(d), _p(r)))
        return r

    def delta_impedance(self, de, pol):
        """extract_delta_impedance_complex (pol 1 or 2) with the H cache of the
        last impedance call."""
        a = np.ascontiguousarray(de, np.complex128)
        z = np.zeros((self.ns, 4), np.complex128)
        _chk(lib().orc_delta_impedance(self.handle, _p(a), int(pol), _p(z)))
        return z

    def solve_frequency(self, freq):
        """Forward solve of both polarisations + impedance, as solve_frequency."""
        self.assemble(2 * np.pi * freq)
        e1, r1 = self.solve_pol(0)
        e2, r2 = self.solve_pol(1)
        z = self.impedance(e1, e2)
        return e1, e2, z, (r1, r2)


def jtj_matvec(model, freqs, z_err, x, inner_tol=1e-4, stale_h=False, fwd=None, sigma_bg=None,
               stations=None):
    """y = J^T W^2 J x (x = delta ln sigma per cell) by the loop of
    ImplicitJtJOperatorFV::Mult (implicit_jtj_fv.cpp:78-192) on the oracle:
    background fields per frequency (compute_single_frequency), then per
    frequency and polarisation p = 1, 2: the perturbation right-hand side with
    E_bg of p, the interleaved solve at inner_tol, delta Z of p, W^2 delta Z
    (w = 1 / error, data with error <= 0 dropped), the adjoint right-hand
    side, the adjoint solve of p's right-hand side at inner_tol, and the
    sensitivity with E_bg of p.  The station H of each frequency comes from
    that frequency's background fields; stale_h=True keeps instead the H of
    the last background frequency for every frequency, as the reference's
    loop does (its station cache is refreshed only by an impedance
    extraction, implicit_jtj_fv.cpp:37-55, 117-119).  Also returns delta Z
    [nf, 2, stations, 4].  The regularisation term of Mult is not formed."""
    from types import SimpleNamespace
    fwd = fwd or model.fwd_params
    inner = SimpleNamespace(**{**fwd.__dict__, "bicgstab_tol": inner_tol})
    o = Oracle(model, fwd=fwd, sigma_bg=sigma_bg)
    oi = Oracle(model, fwd=inner, sigma_bg=sigma_bg)
    sig = model.array("sigma")
    for q in (o, oi):
        q.set_sigma(sig)
        q.find_stations(stations)
    err = np.asarray(z_err)
    w2 = np.where(err > 0, 1.0 / np.where(err > 0, err, 1.0), 0.0) ** 2
    ds = sig * np.asarray(x, np.float64)
    bg = [o.solve_frequency(f)[:2] for f in freqs]
    y = np.zeros(model.nc)
    dz_all = np.zeros((len(freqs), 2, len(stations), 4), np.complex128)
    last = len(freqs) - 1
    for fi, f in enumerate(freqs):
        oi.assemble(2 * np.pi * f)
        # H cache: this frequency's background, or the last one's.
        hf = last if stale_h else fi
        if hf == fi:
            oi.impedance(*bg[fi])
        else:
            oi.assemble(2 * np.pi * freqs[hf])
            oi.impedance(*bg[hf])
            oi.assemble(2 * np.pi * f)
        for p in (1, 2):
            eb = bg[fi][p - 1]
            prhs = oi.perturbation_rhs(eb, ds)
            de, _ = oi.adjoint(prhs)
            dz = oi.delta_impedance(de, p)
            dz_all[fi, p - 1] = dz
            r1, r2 = oi.adjoint_rhs(w2[fi] * dz)
            lam, _ = oi.adjoint(r1 if p == 1 else r2)
            y += oi.sensitivity(eb, lam)
    return y, dz_all


def bicgstab_csr(ptr, col, val, b, x0=None, tol=1e-7, maxiter=2000, ssor=True):
    n = len(ptr) - 1
    ptr = np.ascontiguousarray(ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    x = np.zeros(n, np.complex128) if x0 is None else np.ascontiguousarray(x0, np.complex128).copy()
    r = OrcResult()
    _chk(lib().orc_bicgstab_csr(n, _p(ptr), _p(col), _p(val), _p(b), _p(x), tol, maxiter,
                                int(ssor), C.byref(r)))
    return x, r.as_dict()


def ssor_csr(ptr, col, val, r):
    n = len(ptr) - 1
    ptr = np.ascontiguousarray(ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.complex128)
    r = np.ascontiguousarray(r, np.complex128)
    z = np.zeros(n, np.complex128)
    _chk(lib().orc_ssor_csr(n, _p(ptr), _p(col), _p(val), _p(r), _p(z)))
    return z
