"""ctypes bindings of ``lib/libm3dmt.so`` (include/m3d_device.h, include/m3d_host.h).

The library is loaded from the package's ``lib/`` folder; nothing here falls
back to another implementation: a missing or stale library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libm3dmt.so")

MODE_INNER, MODE_FORWARD, MODE_SYSTEM, MODE_INTERLEAVED = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1

_lib: Optional[C.CDLL] = None


class NativeError(RuntimeError):
    pass


class MeshParamsC(C.Structure):
    _fields_ = [
        ("domain", C.c_double * 6),
        ("min_level", C.c_int), ("max_level", C.c_int),
        ("station_refine_radius", C.c_double),
        ("station_refine_level", C.c_int),
        ("sigma_bg", C.c_double),
        ("max_interest_depth", C.c_double),
        ("air_surface_thickness", C.c_double),
        ("air_mid_thickness", C.c_double),
        ("air_upper_thickness", C.c_double),
        ("transition_padding", C.c_double),
        ("station_bbox", C.c_double * 4),
        ("n_regions", C.c_int),
        ("regions", C.POINTER(C.c_double)),
        ("flat_terrain", C.c_int),
    ]


class ForwardParamsC(C.Structure):
    _fields_ = [
        ("bicgstab_tol", C.c_double),
        ("bicgstab_maxiter", C.c_int),
        ("air_z_threshold", C.c_double),
        ("div_correction", C.c_int),
        ("div_corr_iters", C.c_int),
        ("div_corr_tol", C.c_double),
        ("divcorr_outer_max", C.c_int),
        ("divcorr_iter_per_dc", C.c_int),
    ]


class SolveParamsC(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("maxiter", C.c_int),
        ("div_correction", C.c_int),
        ("div_corr_iters", C.c_int),
        ("div_corr_tol", C.c_double),
        ("divcorr_outer_max", C.c_int),
        ("divcorr_iter_per_dc", C.c_int),
        ("precision", C.c_int),
        ("inner_tol", C.c_double),
        ("use_initial_guess", C.c_int),
        ("interleaved_forward", C.c_int),
        ("divcorr_pcg_tol", C.c_double),
    ]


class SolveResultC(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("residual", C.c_double),
        ("true_residual", C.c_double),
        ("converged", C.c_int),
        ("diverged", C.c_int),
        ("info", C.c_int),
        ("outer", C.c_int),
        ("guess_used", C.c_int),
        ("guess_residual", C.c_double),
        ("stagnated", C.c_int),
    ]


class OperatorDescC(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("rowptr", C.POINTER(C.c_int)),
        ("colidx", C.POINTER(C.c_int)),
        ("values", C.POINTER(C.c_double)),
        ("diag_s", C.POINTER(C.c_double)),
        ("constrained", C.POINTER(C.c_ubyte)),
        ("bc", C.POINTER(C.c_ubyte)),
        ("n_hanging", C.c_int),
        ("hanging", C.POINTER(C.c_int)),
    ]


class DivcorrDescC(C.Structure):
    _fields_ = [
        ("nn", C.c_int),
        ("g_col", C.POINTER(C.c_int)),
        ("g_val", C.POINTER(C.c_double)),
        ("me_sigma", C.POINTER(C.c_double)),
        ("bc_edge", C.POINTER(C.c_ubyte)),
        ("dead_edge", C.POINTER(C.c_ubyte)),
        ("bc_node", C.POINTER(C.c_ubyte)),
        ("L_rowptr", C.POINTER(C.c_int)),
        ("L_colidx", C.POINTER(C.c_int)),
        ("L_values", C.POINTER(C.c_double)),
        ("L_diag", C.POINTER(C.c_double)),
    ]


class LaunchShapeC(C.Structure):
    _fields_ = [
        ("sweep_sub", C.c_int),
        ("sweep_cap", C.c_int),
        ("row_cap", C.c_int),
        ("sweep_chunk", C.c_int),
        ("packed", C.c_int),
        ("wave", C.c_int),
        ("tiled", C.c_int),
    ]


class StationDescC(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("ns", C.c_int), ("max_nf", C.c_int),
        ("valid", C.c_void_p), ("edge_ptr", C.c_void_p), ("edge", C.c_void_p),
        ("role", C.c_void_p), ("cell_size", C.c_void_p), ("nu", C.c_int),
        ("uedge", C.c_void_p), ("contrib_ptr", C.c_void_p), ("contrib", C.c_void_p),
        ("contrib_we", C.c_void_p),
    ]


# Every symbol the two public headers declare.
DEVICE_SYMBOLS = [
    "m3d_context_create", "m3d_context_destroy", "m3d_last_error", "m3d_operator_create",
    "m3d_operator_destroy", "m3d_operator_set_mass", "m3d_operator_set_divcorr", "m3d_solve_batch",
    "m3d_solve_batch_device", "m3d_divcorr", "m3d_spmv", "m3d_ssor_apply",
    "m3d_operator_permutation", "m3d_operator_info", "m3d_operator_format", "m3d_operator_device_bytes", "m3d_operator_trim", "m3d_context_stream", "m3d_bench_kernel",
    "m3d_operator_set_sensitivity", "m3d_sensitivity_device", "m3d_operator_get_shape",
    "m3d_operator_set_shape", "m3d_operator_autotune", "m3d_operator_set_mass_structure",
    "m3d_operator_update_sigma", "m3d_operator_update_sigma_device", "m3d_perturbation_rhs_device",
    "m3d_stations_create", "m3d_stations_destroy", "m3d_impedance_device",
    "m3d_weighted_residual_device", "m3d_adjoint_rhs_device", "m3d_sensitivity_device_noalloc",
    "m3d_operator_set_context", "m3d_stations_set_context",
]
HOST_SYMBOLS = [
    "m3dh_last_error", "m3dh_model_create", "m3dh_model_destroy", "m3dh_get_array",
    "m3dh_get_count", "m3dh_set_sigma", "m3dh_find_stations", "m3dh_primary_field",
    "m3dh_bc_values", "m3dh_rhs_total", "m3dh_phi0", "m3dh_extract_impedance", "m3dh_adjoint_rhs",
    "m3dh_sensitivity", "m3dh_solve_frequencies", "m3dh_adjoint_solve", "m3dh_device_operator",
    "m3dh_update_sigma_device", "m3dh_station_edges", "m3dh_extract_impedance_compact",
    "m3dh_adjoint_rhs_compact", "m3dh_perturbation_rhs", "m3dh_delta_impedance_compact",
    "m3dh_edi_read", "m3dh_edi_write", "m3dh_edi_load_directory", "m3dh_edi_save_directory",
]


def lib() -> C.CDLL:
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(f"{LIB_PATH} is missing: run maple3dmt_amd.build.build() first")
    # The ROCm wheel of torch ships its own HIP and HSA runtimes and asks for
    # them by unversioned names.  Loaded after this library, it would bring a
    # second HSA runtime into the process, which then finds no GPU.  Loaded
    # first, its libamdhip64.so.7 also satisfies this library, so the process
    # has one runtime and device pointers pass freely between the two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i, d, p = C.c_void_p, C.c_int, C.c_double, C.POINTER
    L.m3d_last_error.restype = C.c_char_p
    L.m3dh_last_error.restype = C.c_char_p
    L.m3d_context_create.argtypes = [i, p(vp)]
    L.m3d_context_destroy.argtypes = [vp]
    L.m3d_context_stream.argtypes = [vp]
    L.m3d_context_stream.restype = vp
    L.m3d_operator_create.argtypes = [vp, p(OperatorDescC), p(vp)]
    L.m3d_operator_destroy.argtypes = [vp]
    L.m3d_operator_set_context.argtypes = [vp, vp]
    L.m3d_operator_set_mass.argtypes = [vp, vp]
    L.m3d_operator_set_divcorr.argtypes = [vp, p(DivcorrDescC)]
    L.m3d_solve_batch.argtypes = [vp, i, i, vp, vp, vp, vp, p(SolveParamsC), p(SolveResultC)]
    L.m3d_solve_batch_device.argtypes = [vp, i, i, vp, vp, vp, vp, p(SolveParamsC), p(SolveResultC)]
    L.m3d_divcorr.argtypes = [vp, i, vp, vp, vp, p(SolveParamsC)]
    L.m3d_spmv.argtypes = [vp, i, vp, vp, vp]
    L.m3d_ssor_apply.argtypes = [vp, i, vp, vp, vp]
    L.m3d_operator_permutation.argtypes = [vp, vp, vp]
    L.m3d_operator_info.argtypes = [vp, p(i), p(i), p(C.c_longlong), p(C.c_longlong)]
    L.m3d_operator_format.argtypes = [vp, p(i), p(i)]
    L.m3d_operator_device_bytes.argtypes = [vp, p(C.c_longlong), p(C.c_longlong)]
    L.m3d_operator_trim.argtypes = [vp]
    L.m3d_bench_kernel.argtypes = [vp, i, i, vp, i, p(d), p(d), p(d)]
    L.m3d_sensitivity_device.argtypes = [vp, i, vp, vp, vp, vp]
    L.m3d_sensitivity_device_noalloc.argtypes = [vp, i, vp, vp, vp, vp]
    L.m3d_perturbation_rhs_device.argtypes = [vp, i, vp, vp, vp, vp]
    L.m3d_operator_get_shape.argtypes = [vp, i, i, p(LaunchShapeC)]
    L.m3d_operator_set_shape.argtypes = [vp, i, i, p(LaunchShapeC)]
    L.m3d_operator_autotune.argtypes = [vp, i, vp, i, vp]
    L.m3d_stations_create.argtypes = [vp, p(StationDescC), p(vp)]
    L.m3d_stations_destroy.argtypes = [vp]
    L.m3d_stations_set_context.argtypes = [vp, vp]
    L.m3d_impedance_device.argtypes = [vp, i, vp, vp, vp, vp]
    L.m3d_weighted_residual_device.argtypes = [vp, i, vp, vp, vp, vp, vp]
    L.m3d_adjoint_rhs_device.argtypes = [vp, i, vp, vp]
    L.m3dh_model_create.argtypes = [p(MeshParamsC), vp, i, vp, i, p(ForwardParamsC), vp, p(vp)]
    L.m3dh_model_destroy.argtypes = [vp]
    L.m3dh_get_array.argtypes = [vp, C.c_char_p, p(i), p(C.c_longlong), p(vp)]
    L.m3dh_get_count.argtypes = [vp, C.c_char_p]
    L.m3dh_get_count.restype = C.c_longlong
    L.m3dh_set_sigma.argtypes = [vp, vp]
    L.m3dh_update_sigma_device.argtypes = [vp, vp]
    L.m3dh_station_edges.argtypes = [vp, p(i), p(C.POINTER(C.c_int))]
    L.m3dh_extract_impedance_compact.argtypes = [vp, d, vp, vp, vp]
    L.m3dh_adjoint_rhs_compact.argtypes = [vp, d, vp, vp, vp]
    L.m3dh_perturbation_rhs.argtypes = [vp, d, vp, vp, vp]
    L.m3dh_delta_impedance_compact.argtypes = [vp, vp, i, vp]
    L.m3d_operator_update_sigma.argtypes = [vp, vp]
    L.m3d_operator_update_sigma_device.argtypes = [vp, vp]
    L.m3dh_find_stations.argtypes = [vp, vp, i]
    L.m3dh_primary_field.argtypes = [vp, d, i, vp]
    L.m3dh_bc_values.argtypes = [vp, d, i, vp]
    L.m3dh_rhs_total.argtypes = [vp, vp, vp]
    L.m3dh_phi0.argtypes = [vp, d, vp, vp]
    L.m3dh_extract_impedance.argtypes = [vp, d, vp, vp, vp]
    L.m3dh_adjoint_rhs.argtypes = [vp, d, vp, vp, vp]
    L.m3dh_sensitivity.argtypes = [vp, d, vp, vp, vp]
    L.m3dh_solve_frequencies.argtypes = [vp, vp, i, p(SolveParamsC), vp, vp, p(SolveResultC)]
    L.m3dh_adjoint_solve.argtypes = [vp, vp, i, vp, p(SolveParamsC), vp, p(SolveResultC)]
    L.m3dh_device_operator.argtypes = [vp]
    L.m3dh_device_operator.restype = vp
    _lib = L
    return L


def _check_dev(status: int) -> None:
    if status != 0:
        raise NativeError(lib().m3d_last_error().decode())


def _check_host(status: int) -> None:
    if status != 0:
        raise NativeError(lib().m3dh_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _as_c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


# --------------------------------------------------------------------------
# Parameters
# --------------------------------------------------------------------------
@dataclass
class MeshParams:
    """RefinementParams (octree_mesh.h:37-90) with the reference defaults."""
    domain: Sequence[float]
    min_level: int = 3
    max_level: int = 8
    station_refine_radius: float = 5000.0
    station_refine_level: int = 7
    sigma_bg: float = 0.01
    max_interest_depth: float = 0.0
    air_surface_thickness: float = 2000.0
    air_mid_thickness: float = 8000.0
    air_upper_thickness: float = 20000.0
    transition_padding: float = 0.0
    station_bbox: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    regions: Sequence[Sequence[float]] = field(default_factory=list)
    flat_terrain: bool = True


@dataclass
class ForwardParams:
    """ForwardParamsFV (forward_solver_fv.h:31-69) with the reference defaults."""
    bicgstab_tol: float = 1e-7
    bicgstab_maxiter: int = 2000
    air_z_threshold: float = 1.0
    div_correction: bool = True
    div_corr_iters: int = 10
    div_corr_tol: float = 2e-3
    divcorr_outer_max: int = 20
    divcorr_iter_per_dc: int = 500

    def solve_params(self, precision: int = PREC_FP64, inner_tol: float = 0.0,
                     tol: Optional[float] = None) -> "SolveParams":
        return SolveParams(tol=self.bicgstab_tol if tol is None else tol,
                           maxiter=self.bicgstab_maxiter, div_correction=self.div_correction,
                           div_corr_iters=self.div_corr_iters, div_corr_tol=self.div_corr_tol,
                           divcorr_outer_max=self.divcorr_outer_max,
                           divcorr_iter_per_dc=self.divcorr_iter_per_dc, precision=precision,
                           inner_tol=inner_tol)

    def fast_params(self, precision: int = PREC_MIXED, dc_interval: int = 100,
                    tol: Optional[float] = None) -> "SolveParams":
        """The batched fast path: E_sec by bursts of dc_interval BiCGStab
        iterations (complex64 inner solves in FP64 refinement by default), each
        followed by DivCorr whose PCG stops at 1e-4 relative instead of the
        reference's 1e-10 (the next burst corrects what a looser potential
        leaves; measured on the 1.45 M-edge random model the bursts per system
        do not change while the DivCorr work drops ~2.5x).  The result meets
        the same FP64 true-residual tolerance."""
        prm = self.solve_params(precision=precision, inner_tol=1e-3 if precision else 0.0, tol=tol)
        prm.interleaved_forward = 1
        prm.divcorr_iter_per_dc = dc_interval
        prm.divcorr_outer_max = 400
        prm.divcorr_pcg_tol = 1e-4
        return prm


@dataclass
class SolveParams:
    tol: float = 1e-7
    maxiter: int = 2000
    div_correction: bool = True
    div_corr_iters: int = 10
    div_corr_tol: float = 2e-3
    divcorr_outer_max: int = 20
    divcorr_iter_per_dc: int = 500
    precision: int = PREC_FP64
    inner_tol: float = 0.0
    use_initial_guess: int = 0   # 0 zero start, 1 guess as given, 2 or 3 least-squares scaled guess (kept below 0.5 or 1 relative residual)
    interleaved_forward: int = 0  # FORWARD E_sec by interleaved BiCGStab + DivCorr bursts
    divcorr_pcg_tol: float = 0.0  # DivCorr PCG relative stop; 0 = the reference's 1e-10

    def c(self) -> SolveParamsC:
        return SolveParamsC(self.tol, self.maxiter, int(self.div_correction), self.div_corr_iters,
                            self.div_corr_tol, self.divcorr_outer_max, self.divcorr_iter_per_dc,
                            self.precision, self.inner_tol, int(self.use_initial_guess),
                            int(self.interleaved_forward), float(self.divcorr_pcg_tol))


def results_to_dicts(res) -> list:
    return [dict(iterations=r.iterations, residual=r.residual, true_residual=r.true_residual,
                 converged=bool(r.converged), diverged=bool(r.diverged), info=r.info,
                 outer=r.outer, guess_used=bool(r.guess_used),
                 guess_residual=r.guess_residual, stagnated=bool(r.stagnated)) for r in res]


# --------------------------------------------------------------------------
# Device context and operator
# --------------------------------------------------------------------------
class Context:
    """One device context (m3d_context_create)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check_dev(lib().m3d_context_create(device, C.byref(h)))
        self.handle = h

    @property
    def stream(self) -> int:
        return lib().m3d_context_stream(self.handle) or 0

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().m3d_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceOperator:
    """Thin handle over an m3d_operator (owned or borrowed)."""

    def __init__(self, handle, owner=None, n: int = 0, nn: int = 0):
        self.handle = handle
        self._owner = owner
        self.n = n
        self.nn = nn

    def info(self) -> dict:
        npad, nc = C.c_int(), C.c_int()
        st, nz = C.c_longlong(), C.c_longlong()
        _check_dev(lib().m3d_operator_info(self.handle, C.byref(npad), C.byref(nc), C.byref(st),
                                           C.byref(nz)))
        tab, cb = C.c_int(), C.c_int()
        _check_dev(lib().m3d_operator_format(self.handle, C.byref(tab), C.byref(cb)))
        return dict(n_padded=npad.value, ncolours=nc.value, stored=st.value, nnz=nz.value,
                    value_table=tab.value, column_bits=cb.value)

    def device_bytes(self) -> dict:
        """Device memory of the operator's arrays and of its solve workspaces."""
        ob, wb = C.c_longlong(), C.c_longlong()
        _check_dev(lib().m3d_operator_device_bytes(self.handle, C.byref(ob), C.byref(wb)))
        return {"operator": ob.value, "workspaces": wb.value}

    def trim(self) -> None:
        """Free the solve workspaces (the next solve allocates them again)."""
        _check_dev(lib().m3d_operator_trim(self.handle))

    def permutation(self):
        inf = self.info()
        perm = np.zeros(inf["n_padded"], np.int32)
        cs = np.zeros(inf["ncolours"] + 1, np.int32)
        _check_dev(lib().m3d_operator_permutation(self.handle, _ptr(perm), _ptr(cs)))
        return perm, cs

    def bench_kernel(self, kind: int, omegas, reps: int = 20, stored: bool = False):
        """(avg ms per launch, algorithmic bytes per launch) of kind 0 = SpMV,
        1 = SSOR apply, 2 = one FP64 BiCGStab iteration, 3 = one complex64
        BiCGStab iteration, 4 = one backward-sweep launch (mean over the
        colours), timed with HIP events on the library stream; the bytes
        count a coupling as its value and int32 column.  With stored, a third
        value: the bytes with the stored format's bytes per coupling."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        ms, by, sb = C.c_double(), C.c_double(), C.c_double()
        _check_dev(lib().m3d_bench_kernel(self.handle, kind, om.shape[0], _ptr(om), reps,
                                          C.byref(ms), C.byref(by), C.byref(sb)))
        return (ms.value, by.value, sb.value) if stored else (ms.value, by.value)

    def update_sigma_device(self, d_sigma: int) -> None:
        """m3d_operator_update_sigma_device: sigma [cells] in device memory."""
        _check_dev(lib().m3d_operator_update_sigma_device(self.handle, C.c_void_p(d_sigma)))

    def get_shape(self, precision: int, width: int) -> dict:
        """Launch shape of the row kernels at one precision and batch width."""
        sh = LaunchShapeC()
        _check_dev(lib().m3d_operator_get_shape(self.handle, precision, width, C.byref(sh)))
        return {"sweep_sub": sh.sweep_sub, "sweep_cap": sh.sweep_cap, "row_cap": sh.row_cap,
                "sweep_chunk": sh.sweep_chunk, "packed": sh.packed, "wave": sh.wave, "tiled": sh.tiled}

    def set_shape(self, precision: int, width: int, sweep_sub: int, sweep_cap: int, row_cap: int,
                  sweep_chunk: int = 1, packed: Optional[int] = None, wave: Optional[int] = None,
                  tiled: Optional[int] = None) -> None:
        """packed=None / wave=None / tiled=None keep the width's current choice."""
        cur = self.get_shape(precision, width)
        packed = cur["packed"] if packed is None else packed
        wave = cur["wave"] if wave is None else wave
        tiled = cur["tiled"] if tiled is None else tiled
        sh = LaunchShapeC(sweep_sub, sweep_cap, row_cap, sweep_chunk, packed, wave, tiled)
        _check_dev(lib().m3d_operator_set_shape(self.handle, precision, width, C.byref(sh)))

    def autotune(self, omegas, widths=(1, 2, 4, 8, 16), reps: int = 16) -> dict:
        """Times candidate launch shapes of one BiCGStab iteration per width
        and precision on this operator and keeps the fastest; returns
        {(precision, width): (ms per iteration, shape)}."""
        mask = 0
        for w in widths:
            mask |= 1 << (int(w).bit_length() - 1)
        om = np.ascontiguousarray(np.resize(np.asarray(omegas, dtype=np.float64), 32))
        best = np.zeros(12)
        _check_dev(lib().m3d_operator_autotune(self.handle, mask, _ptr(om), reps, _ptr(best)))
        out = {}
        for prec in (PREC_FP64, PREC_MIXED):
            for w in widths:
                slot = int(w).bit_length() - 1
                out[(prec, int(w))] = (float(best[prec * 6 + slot]), self.get_shape(prec, int(w)))
        return out

    def perturbation_rhs_device(self, omegas, d_ebg: int, d_dsigma: int, d_rhs: int) -> None:
        """m3d_perturbation_rhs_device: device pointers, E_bg and rhs complex128
        [nsys][n], dsigma float64 [cells]."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        _check_dev(lib().m3d_perturbation_rhs_device(self.handle, om.shape[0], _ptr(om), C.c_void_p(d_ebg),
                                                     C.c_void_p(d_dsigma), C.c_void_p(d_rhs)))

    def sensitivity_device(self, omegas, d_e: int, d_lambda: int, d_grad: int) -> None:
        """d_grad[c] = sum over systems of Re(i w sigma_c sum_e lambda E w_e),
        all device pointers (E, lambda complex128 [nsys][n], grad float64 [nc])."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        _check_dev(lib().m3d_sensitivity_device(self.handle, om.shape[0], _ptr(om), C.c_void_p(d_e),
                                                C.c_void_p(d_lambda), C.c_void_p(d_grad)))

    def sensitivity_device_noalloc(self, omegas, d_e: int, d_lambda: int, d_grad: int) -> None:
        """sensitivity_device with no device allocation: the frequencies are
        staged in the library's static per-GPU array (at most 4096)."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        _check_dev(lib().m3d_sensitivity_device_noalloc(self.handle, om.shape[0], _ptr(om), C.c_void_p(d_e),
                                                        C.c_void_p(d_lambda), C.c_void_p(d_grad)))

    def spmv(self, omegas, x) -> np.ndarray:
        x = _as_c128(np.atleast_2d(x))
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        y = np.zeros_like(x)
        _check_dev(lib().m3d_spmv(self.handle, x.shape[0], _ptr(om), _ptr(x), _ptr(y)))
        return y

    def ssor_apply(self, omegas, r) -> np.ndarray:
        r = _as_c128(np.atleast_2d(r))
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        z = np.zeros_like(r)
        _check_dev(lib().m3d_ssor_apply(self.handle, r.shape[0], _ptr(om), _ptr(r), _ptr(z)))
        return z

    def divcorr(self, omegas, E, phi0=None, params: Optional[SolveParams] = None) -> np.ndarray:
        E = _as_c128(np.atleast_2d(E)).copy()
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        p0 = None if phi0 is None else _as_c128(np.atleast_2d(phi0))
        prm = (params or SolveParams()).c()
        _check_dev(lib().m3d_divcorr(self.handle, E.shape[0], _ptr(om), _ptr(E),
                                     None if p0 is None else _ptr(p0), C.byref(prm)))
        return E

    def solve(self, mode: int, omegas, rhs, aux=None, x=None,
              params: Optional[SolveParams] = None):
        rhs = _as_c128(np.atleast_2d(rhs))
        ns = rhs.shape[0]
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        assert om.shape == (ns,)
        xx = np.zeros_like(rhs) if x is None else _as_c128(np.atleast_2d(x)).copy()
        a = None if aux is None else _as_c128(np.atleast_2d(aux))
        prm = (params or SolveParams()).c()
        res = (SolveResultC * ns)()
        _check_dev(lib().m3d_solve_batch(self.handle, ns, mode, _ptr(om), _ptr(rhs),
                                         None if a is None else _ptr(a), _ptr(xx), C.byref(prm),
                                         res))
        return xx, results_to_dicts(res)

    def solve_device(self, mode: int, omegas, d_rhs: int, d_aux: int, d_x: int, nsys: int,
                     params: Optional[SolveParams] = None):
        """Solve with device pointers (e.g. torch tensor .data_ptr())."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        prm = (params or SolveParams()).c()
        res = (SolveResultC * nsys)()
        _check_dev(lib().m3d_solve_batch_device(self.handle, nsys, mode, _ptr(om),
                                                C.c_void_p(d_rhs),
                                                C.c_void_p(d_aux) if d_aux else None,
                                                C.c_void_p(d_x), C.byref(prm), res))
        return results_to_dicts(res)


# --------------------------------------------------------------------------
# Host model (mesh + grid + operators + ForwardSolverFV-shaped driver)
# --------------------------------------------------------------------------
_DT = {0: np.float64, 1: np.int32, 2: np.uint8, 3: np.complex128}


class DeviceStations:
    """The station side of the gradient on the device (m3d_stations_*): the
    impedance of device fields, the weighted residual and misfit, and the
    adjoint right-hand side, with every buffer in HBM (device pointers as
    ints, e.g. torch tensors' data_ptr()).  Built from a Model's station
    table for at most max_nf frequencies per call."""

    def __init__(self, model: "Model", ctx: Context, max_nf: int):
        keep = {k: model.array("st_" + k) for k in
                ("valid", "edge_ptr", "edge", "role", "cell_size", "uedge", "contrib_ptr", "contrib",
                 "contrib_we")}
        keep = {k: np.ascontiguousarray(v) for k, v in keep.items()}
        d = StationDescC()
        d.n, d.ns, d.max_nf = model.n, len(keep["valid"]), int(max_nf)
        d.nu = len(keep["uedge"])
        for k, v in keep.items():
            setattr(d, k, _ptr(v) if v.size else None)
        h = C.c_void_p()
        _check_dev(lib().m3d_stations_create(ctx.handle, C.byref(d), C.byref(h)))
        self.handle = h
        self.ns, self.max_nf, self.n = d.ns, d.max_nf, d.n
        self.sigma = np.ascontiguousarray(model.array("st_sigma"), dtype=np.float64)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().m3d_stations_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def impedance(self, omegas, d_e: int, d_z: int) -> None:
        """Z [nf][ns][4] of the fields [2 nf][n] (m3d_impedance_device)."""
        om = np.ascontiguousarray(omegas, dtype=np.float64).reshape(-1)
        _check_dev(lib().m3d_impedance_device(self.handle, len(om), _ptr(om),
                                              _ptr(self.sigma) if self.ns else None, d_e, d_z))

    def weighted_residual(self, nf: int, d_obs: int, d_err: int, d_weight: Optional[int], d_wr: int,
                          d_phi: int) -> None:
        _check_dev(lib().m3d_weighted_residual_device(self.handle, nf, d_obs, d_err, d_weight or None,
                                                      d_wr, d_phi))

    def adjoint_rhs(self, nf: int, d_wr: int, d_rhs: int) -> None:
        _check_dev(lib().m3d_adjoint_rhs_device(self.handle, nf, d_wr, d_rhs))


class Model:
    """Octree mesh, staggered grid, operators and host solve data of one
    conductivity model; with a Context it also owns the device operator."""

    def __init__(self, mesh: MeshParams, stations_xyz, freqs, fwd: Optional[ForwardParams] = None,
                 ctx: Optional[Context] = None):
        self.mesh_params = mesh
        self.fwd_params = fwd or ForwardParams()
        st = np.ascontiguousarray(np.asarray(stations_xyz, dtype=np.float64).reshape(-1, 3))
        fr = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64).reshape(-1))
        regs = np.ascontiguousarray(np.asarray(mesh.regions, dtype=np.float64).reshape(-1))
        mp = MeshParamsC()
        mp.domain[:] = list(mesh.domain)
        mp.min_level, mp.max_level = mesh.min_level, mesh.max_level
        mp.station_refine_radius = mesh.station_refine_radius
        mp.station_refine_level = mesh.station_refine_level
        mp.sigma_bg = mesh.sigma_bg
        mp.max_interest_depth = mesh.max_interest_depth
        mp.air_surface_thickness = mesh.air_surface_thickness
        mp.air_mid_thickness = mesh.air_mid_thickness
        mp.air_upper_thickness = mesh.air_upper_thickness
        mp.transition_padding = mesh.transition_padding
        mp.station_bbox[:] = list(mesh.station_bbox)
        mp.n_regions = len(mesh.regions)
        mp.regions = regs.ctypes.data_as(C.POINTER(C.c_double)) if len(regs) else None
        mp.flat_terrain = int(mesh.flat_terrain)
        f = self.fwd_params
        fp = ForwardParamsC(f.bicgstab_tol, f.bicgstab_maxiter, f.air_z_threshold,
                            int(f.div_correction), f.div_corr_iters, f.div_corr_tol,
                            f.divcorr_outer_max, f.divcorr_iter_per_dc)
        h = C.c_void_p()
        _check_host(lib().m3dh_model_create(C.byref(mp), _ptr(st), st.shape[0], _ptr(fr),
                                            fr.shape[0], C.byref(fp),
                                            ctx.handle if ctx is not None else None, C.byref(h)))
        self.handle = h
        self.ctx = ctx
        self.n = int(self.count("edges"))
        self.nc = int(self.count("cells"))
        self.nn = int(self.count("nodes"))
        self._stations = None
        self._dev_stations = None

    def close(self) -> None:
        self._drop_stations()
        if getattr(self, "handle", None):
            lib().m3dh_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- arrays -----------------------------------------------------------
    def array(self, name: str) -> np.ndarray:
        dt, cnt, ptr = C.c_int(), C.c_longlong(), C.c_void_p()
        _check_host(lib().m3dh_get_array(self.handle, name.encode(), C.byref(dt), C.byref(cnt),
                                         C.byref(ptr)))
        dtype = _DT[dt.value]
        if cnt.value == 0:
            return np.zeros(0, dtype)
        buf = (C.c_char * (cnt.value * np.dtype(dtype).itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dtype).copy()

    def count(self, name: str) -> int:
        return int(lib().m3dh_get_count(self.handle, name.encode()))

    @property
    def device(self) -> DeviceOperator:
        h = lib().m3dh_device_operator(self.handle)
        if not h:
            raise NativeError("model has no device operator (created without a Context)")
        return DeviceOperator(C.c_void_p(h), owner=self, n=self.n, nn=self.nn)

    # -- host producers ----------------------------------------------------
    def set_sigma(self, sigma) -> None:
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        assert s.shape == (self.nc,)
        _check_host(lib().m3dh_set_sigma(self.handle, _ptr(s)))

    def update_sigma_device(self, sigma) -> None:
        """set_sigma for a line-search trial after one set_sigma: the device
        recomputes Me and the DivCorr Laplacian (the same arrays set_sigma
        would upload); the host keeps the new sigma."""
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        assert s.shape == (self.nc,)
        _check_host(lib().m3dh_update_sigma_device(self.handle, _ptr(s)))

    def device_stations(self, max_nf: int) -> DeviceStations:
        """The device station set of the current stations (kept and reused
        while the stations stay the same and max_nf fits; the station cells'
        conductivity for the 1-D fallback is taken again from the current
        sigma)."""
        if self.ctx is None:
            raise NativeError("model has no device operator (created without a Context)")
        st = self._dev_stations
        if st is None or st.max_nf < max_nf or st.handle is None:
            if st is not None:
                st.close()
            self._dev_stations = st = DeviceStations(self, self.ctx, max_nf)
        else:
            st.sigma = np.ascontiguousarray(self.array("st_sigma"), dtype=np.float64)
        return st

    def set_context(self, ctx: Context) -> None:
        """Moves the device operator and the device station set to ctx (same
        device): waits for the current stream, then every later call runs on
        ctx's stream with the same workspaces, launch shapes and graphs."""
        if self.ctx is None:
            raise NativeError("model has no device operator (created without a Context)")
        if ctx is self.ctx:
            return
        _check_dev(lib().m3d_operator_set_context(self.device.handle, ctx.handle))
        st = self._dev_stations
        if st is not None and st.handle is not None:
            _check_dev(lib().m3d_stations_set_context(st.handle, ctx.handle))
        self.ctx = ctx

    def _drop_stations(self) -> None:
        st = getattr(self, "_dev_stations", None)
        if st is not None:
            st.close()
            self._dev_stations = None

    def find_stations(self, xy) -> np.ndarray:
        a = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
        _check_host(lib().m3dh_find_stations(self.handle, _ptr(a), a.shape[0]))
        self._drop_stations()
        self._stations = a.shape[0]
        return self.array("station_cell")

    def primary_field(self, omega: float, direction: int) -> np.ndarray:
        out = np.zeros(self.n, np.complex128)
        _check_host(lib().m3dh_primary_field(self.handle, omega, direction, _ptr(out)))
        return out

    def bc_values(self, omega: float, direction: int) -> np.ndarray:
        out = np.zeros(self.n, np.complex128)
        _check_host(lib().m3dh_bc_values(self.handle, omega, direction, _ptr(out)))
        return out

    def rhs_total(self, ebc) -> np.ndarray:
        e = _as_c128(ebc)
        out = np.zeros(self.n, np.complex128)
        _check_host(lib().m3dh_rhs_total(self.handle, _ptr(e), _ptr(out)))
        return out

    def phi0(self, omega: float, rhs) -> np.ndarray:
        r = _as_c128(rhs)
        out = np.zeros(self.nn, np.complex128)
        _check_host(lib().m3dh_phi0(self.handle, omega, _ptr(r), _ptr(out)))
        return out

    def extract_impedance(self, omega: float, e1, e2) -> np.ndarray:
        a, b = _as_c128(e1), _as_c128(e2)
        ns = self.count("stations")
        z = np.zeros((ns, 4), np.complex128)
        _check_host(lib().m3dh_extract_impedance(self.handle, omega, _ptr(a), _ptr(b), _ptr(z)))
        return z

    def adjoint_rhs(self, omega: float, wr):
        w = _as_c128(wr)
        r1 = np.zeros(self.n, np.complex128)
        r2 = np.zeros(self.n, np.complex128)
        _check_host(lib().m3dh_adjoint_rhs(self.handle, omega, _ptr(w), _ptr(r1), _ptr(r2)))
        return r1, r2

    def station_edges(self) -> np.ndarray:
        """Sorted edges of the station cells (the compact station arrays'
        index)."""
        n = C.c_int()
        ptr = C.POINTER(C.c_int)()
        _check_host(lib().m3dh_station_edges(self.handle, C.byref(n), C.byref(ptr)))
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy() if n.value else np.zeros(0, np.int32)

    def extract_impedance_compact(self, omega: float, e1c, e2c) -> np.ndarray:
        """extract_impedance from the values on station_edges() only."""
        a, b = _as_c128(e1c), _as_c128(e2c)
        z = np.zeros((self.count("stations"), 4), np.complex128)
        _check_host(lib().m3dh_extract_impedance_compact(self.handle, omega, _ptr(a), _ptr(b), _ptr(z)))
        return z

    def adjoint_rhs_compact(self, omega: float, wr):
        """adjoint_rhs on station_edges() only (the rest of both RHS is 0)."""
        w = _as_c128(wr)
        k = len(self.station_edges())
        r1 = np.zeros(k, np.complex128)
        r2 = np.zeros(k, np.complex128)
        _check_host(lib().m3dh_adjoint_rhs_compact(self.handle, omega, _ptr(w), _ptr(r1), _ptr(r2)))
        return r1, r2

    def perturbation_rhs(self, omega: float, e_bg, dsigma) -> np.ndarray:
        """build_perturbation_rhs: -i w sum_c dMe/dsigma_c dsigma_c E_bg with the
        reference's harmonic-mean L^3 mass derivative (dsigma in sigma space)."""
        e = _as_c128(e_bg)
        ds = np.ascontiguousarray(dsigma, dtype=np.float64)
        assert ds.shape == (self.nc,)
        out = np.zeros(self.n, np.complex128)
        _check_host(lib().m3dh_perturbation_rhs(self.handle, omega, _ptr(e), _ptr(ds), _ptr(out)))
        return out

    def delta_impedance_compact(self, dec, polarization: int) -> np.ndarray:
        """extract_delta_impedance_complex from delta E on station_edges()
        (polarization 1 or 2; the H cache of the last impedance extraction)."""
        a = _as_c128(dec)
        dz = np.zeros((self.count("stations"), 4), np.complex128)
        _check_host(lib().m3dh_delta_impedance_compact(self.handle, _ptr(a), polarization, _ptr(dz)))
        return dz

    def sensitivity(self, omega: float, e_bg, lam) -> np.ndarray:
        a, b = _as_c128(e_bg), _as_c128(lam)
        out = np.zeros(self.nc, np.float64)
        _check_host(lib().m3dh_sensitivity(self.handle, omega, _ptr(a), _ptr(b), _ptr(out)))
        return out

    # -- device solves -----------------------------------------------------
    def solve_frequencies(self, freqs, params: Optional[SolveParams] = None):
        """Forward solve (both polarisations) of every frequency in one batch.
        Returns E [2*nf, n] (row 2f+pol), Z [nf, stations, 4] and results."""
        fr = np.ascontiguousarray(freqs, dtype=np.float64).reshape(-1)
        nf = fr.shape[0]
        ns = self.count("stations")
        e = np.zeros((2 * nf, self.n), np.complex128)
        z = np.zeros((nf, max(ns, 0), 4), np.complex128)
        prm = (params or self.fwd_params.solve_params()).c()
        res = (SolveResultC * (2 * nf))()
        _check_host(lib().m3dh_solve_frequencies(self.handle, _ptr(fr), nf, C.byref(prm), _ptr(e),
                                                 _ptr(z) if ns > 0 else None, res))
        return e, z, results_to_dicts(res)

    def adjoint_solve(self, omegas, rhs, params: Optional[SolveParams] = None):
        r = _as_c128(np.atleast_2d(rhs))
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        x = np.zeros_like(r)
        prm = (params or self.fwd_params.solve_params()).c()
        res = (SolveResultC * r.shape[0])()
        _check_host(lib().m3dh_adjoint_solve(self.handle, _ptr(om), r.shape[0], _ptr(r),
                                             C.byref(prm), _ptr(x), res))
        return x, results_to_dicts(res)
