/* C API of the host side: the p4est-free octree / staggered-grid / operator
 * producer and the ForwardSolverFV-shaped driver that reaches the device
 * through m3d_device.h.  Used by bindings (ctypes) and by C/C++ drivers; a
 * C++ driver can use maple3dmt_amd/csrc/host/forward_solver_gpu.h directly.
 *
 * Reference interfaces mirrored:
 *   m3dh_model_create   OctreeMesh::setup + set_terrain (octree_mesh.cpp:25-78),
 *                       StaggeredGrid::build (staggered_grid.cpp:70-236),
 *                       DiscreteOperators::build (operators.cpp:19-170),
 *                       ForwardSolverFV::setup (forward_solver_fv.cpp:28-47)
 *   m3dh_set_sigma      ForwardSolverFV::set_sigma / update_sigma (:49-58)
 *   m3dh_update_sigma_device  the same update with Me and the DivCorr
 *                       Laplacian recomputed on the device (line-search
 *                       trials, inversion_fv.cpp:427-436)
 *   m3dh_find_stations  find_stations_ (:1393-1485)
 *   m3dh_solve_frequencies   solve_frequency (:84-531) for several frequencies
 *   m3dh_extract_impedance   extract_impedance_ (:1490-1727)
 *   m3dh_adjoint_rhs    build_adjoint_rhs_from_residual (:1804-2013)
 *   m3dh_*_compact      the same two on the station cells' edges only
 *   m3dh_adjoint_solve  adjoint_solve_complex / solve_rhs_complex (:914-940)
 *   m3dh_sensitivity    compute_sensitivity (:989-1059)
 */
#ifndef M3D_HOST_H
#define M3D_HOST_H

#include "m3d_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct m3dh_model m3dh_model;

/* RefinementParams (octree_mesh.h:37-90) plus flat-terrain classification. */
typedef struct {
    double domain[6]; /* x_min, x_max, y_min, y_max, z_min, z_max */
    int min_level, max_level;
    double station_refine_radius;
    int station_refine_level;
    double sigma_bg;
    double max_interest_depth;
    double air_surface_thickness, air_mid_thickness, air_upper_thickness;
    double transition_padding;
    double station_bbox[4]; /* station_x_min, x_max, y_min, y_max */
    int n_regions;
    const double* regions;  /* 8 per region: x_min,x_max,y_min,y_max,z_min,z_max,level,padding */
    int flat_terrain;       /* 1: set_terrain(z = 0) as the reference drivers do */
} m3dh_mesh_params;

/* ForwardParamsFV (forward_solver_fv.h:31-69). */
typedef struct {
    double bicgstab_tol;
    int bicgstab_maxiter;
    double air_z_threshold;
    int div_correction;
    int div_corr_iters;
    double div_corr_tol;
    int divcorr_outer_max;
    int divcorr_iter_per_dc;
} m3dh_forward_params;

const char* m3dh_last_error(void);

/* Build mesh, grid, operators and host flags.  ctx may be NULL (host-only
 * model: no device operator). */
int m3dh_model_create(const m3dh_mesh_params* mesh, const double* stations_xyz, int n_stations,
                      const double* freqs, int n_freqs, const m3dh_forward_params* fwd,
                      m3d_context* ctx, m3dh_model** out);
int m3dh_model_destroy(m3dh_model* m);

/* Named array of the model, e.g. "edge_x", "S_val", "L_ptr" (see
 * host_capi.cpp for the list).  dtype: 0 = float64, 1 = int32, 2 = uint8,
 * 3 = complex128.  The pointer stays valid until the model changes or is
 * destroyed. */
int m3dh_get_array(m3dh_model* m, const char* name, int* dtype, long long* count, const void** data);
/* Scalar counts: "cells", "edges", "faces", "nodes", "hanging_faces",
 * "hanging_edges", "stations", "colours". */
long long m3dh_get_count(m3dh_model* m, const char* name);

int m3dh_set_sigma(m3dh_model* m, const double* sigma);
/* After one m3dh_set_sigma: a new sigma [cells] applied by
 * m3d_operator_update_sigma (device arrays equal to m3dh_set_sigma's). */
int m3dh_update_sigma_device(m3dh_model* m, const double* sigma);
int m3dh_find_stations(m3dh_model* m, const double* xy, int n);

/* Host producers for one polarisation (dir 0 = pol 1 / Ex, 1 = pol 2 / Ey). */
int m3dh_primary_field(m3dh_model* m, double omega, int dir, double* out_c128);
int m3dh_bc_values(m3dh_model* m, double omega, int dir, double* out_c128);
int m3dh_rhs_total(m3dh_model* m, const double* ebc_c128, double* out_c128);
int m3dh_phi0(m3dh_model* m, double omega, const double* rhs_c128, double* out_c128);

/* Z from E1, E2 (n complex each): z_out is [stations][4] complex (Zxx, Zxy,
 * Zyx, Zyy).  The station H cache is kept in the model for m3dh_adjoint_rhs. */
int m3dh_extract_impedance(m3dh_model* m, double omega, const double* e1, const double* e2,
                           double* z_out);
int m3dh_adjoint_rhs(m3dh_model* m, double omega, const double* wr_c128, double* rhs1,
                     double* rhs2);
int m3dh_sensitivity(m3dh_model* m, double omega, const double* e_bg, const double* lambda,
                     double* sens);

/* The same station steps on compact arrays, for fields that stay on the
 * device: m3dh_station_edges gives the sorted edges of the station cells
 * (count values; the pointer stays valid until the next call or
 * m3dh_find_stations); e1c / e2c and rhs1c / rhs2c hold one complex value per
 * listed edge.  Z and the station H cache equal m3dh_extract_impedance's on
 * the full fields, and the right-hand sides equal m3dh_adjoint_rhs's on the
 * listed edges (zero elsewhere). */
int m3dh_station_edges(m3dh_model* m, int* count, const int** edges);

/* GN-CG J*v pieces (ImplicitJtJOperatorFV::Mult, implicit_jtj_fv.cpp:78-192):
 * the perturbation right-hand side -i omega sum_c dMe/dsigma_c dsigma_c E_bg
 * with the reference's harmonic-mean L^3 mass derivative
 * (build_perturbation_rhs, forward_solver_fv.cpp:945-984; dsigma in sigma
 * space, [cells]), and delta Z at the stations from delta E on the station
 * edges (extract_delta_impedance_complex, :2018-2065; polarization 1 or 2;
 * uses the station H cache of the last impedance extraction). */
int m3dh_perturbation_rhs(m3dh_model* m, double omega, const double* e_bg_c128, const double* dsigma,
                          double* rhs_c128);
int m3dh_delta_impedance_compact(m3dh_model* m, const double* dec_c128, int polarization,
                                 double* dz_out);
int m3dh_extract_impedance_compact(m3dh_model* m, double omega, const double* e1c_c128,
                                   const double* e2c_c128, double* z_out);
int m3dh_adjoint_rhs_compact(m3dh_model* m, double omega, const double* wr_c128, double* rhs1c,
                             double* rhs2c);

/* Device solves through the model's device operator. */
int m3dh_solve_frequencies(m3dh_model* m, const double* freqs, int nf, const m3d_solve_params* prm,
                           double* e_out, double* z_out, m3d_solve_result* results);
int m3dh_adjoint_solve(m3dh_model* m, const double* omegas, int nsys, const double* rhs,
                       const m3d_solve_params* prm, double* x_out, m3d_solve_result* results);
/* Device operator of the model (NULL for host-only models). */
m3d_operator* m3dh_device_operator(m3dh_model* m);

/* EDI files (io::read_edi / write_edi / load_edi_directory /
 * save_edi_directory, src/io/edi_io.cpp:156-619): one station per file,
 * impedances in ohm (EDI mV/km/nT times 4 pi 1e-4), errors from the EDI
 * variances, weight 0 for a component the file lacks.  Arrays, for ns
 * stations and nf frequencies (index f * ns + s for the directory calls):
 *   names [ns][name_len] chars, geo [ns][4] = {lat, lon, elev, has_geo},
 *   freqs [nf], z [..][4] complex (Zxx, Zxy, Zyx, Zyy), z_err [..][4],
 *   z_w [..][4], t [..][2] complex (Tx, Ty), t_err [..][2], t_w [..][2].
 * Reading calls fill the counts always and the arrays that are not NULL, so
 * a first call with NULL arrays sizes the second.  Directory loading takes
 * the union of the stations' frequencies, highest first, with weight 0 where
 * a station lacks a frequency; t, t_err, t_w may be NULL on writes (no
 * tipper). */
int m3dh_edi_read(const char* path, int* nfreq, char* name, int name_len, double* geo, double* freqs,
                  double* z, double* z_err, double* z_w, double* t, double* t_err, double* t_w);
int m3dh_edi_write(const char* path, const char* name, const double* geo, int nfreq, const double* freqs,
                   const double* z, const double* z_err, const double* z_w, const double* t,
                   const double* t_err, const double* t_w);
int m3dh_edi_load_directory(const char* dir, int* ns, int* nf, char* names, int name_len, double* geo,
                            double* freqs, double* z, double* z_err, double* z_w, double* t,
                            double* t_err, double* t_w);
int m3dh_edi_save_directory(const char* dir, int ns, int nf, const char* names, int name_len,
                            const double* geo, const double* freqs, const double* z, const double* z_err,
                            const double* z_w, const double* t, const double* t_err, const double* t_w);

#ifdef __cplusplus
}
#endif

#endif /* M3D_HOST_H */
