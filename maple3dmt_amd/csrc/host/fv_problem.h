// Host-side pieces of the octree FV forward / adjoint solve that sit either
// side of the device solve: constraint flags, Dirichlet values, right-hand
// sides, the frequency-independent device operator, DivCorr setup, station
// mapping, impedance extraction, adjoint right-hand side and sensitivity.
//
// These restate the host work of ForwardSolverFV
// (src/forward/forward_solver_fv.cpp): BC / dead-edge flags and identity
// elimination (:111-235), right-hand side build (:244-255), primary field and
// scattered right-hand side (:121-136, :316-334), DivCorr flags and Laplacian
// (:546-610), phi0 (:893-908), station mapping (:1393-1485), impedance
// (:1490-1727), adjoint right-hand side (:1804-2013), sensitivity
// (:989-1059), perturbation right-hand side (:945-984) and delta impedance
// (:2018-2065).
#pragma once

#include <array>
#include <complex>
#include <vector>

#include "operators.h"
#include "staggered_grid.h"

namespace m3d {

using cplx = std::complex<double>;

// Mirrors ForwardParamsFV (include/maple3dmt/forward/forward_solver_fv.h:31-69).
struct ForwardParams {
    double bicgstab_tol = 1e-7;
    int bicgstab_maxiter = 2000;
    int print_level = 0;
    bool scattered_field = true;
    double air_z_threshold = 1.0;
    int air_bc_iterations = 1;
    bool div_correction = true;
    int div_corr_iters = 10;
    double div_corr_tol = 2e-3;
    int divcorr_outer_max = 20;
    int divcorr_iter_per_dc = 500;
};

// Per-station H, Z and cell sizes cached by the impedance extraction
// (forward_solver_fv.h:216-224).
struct StationHCache {
    cplx Hx1, Hy1, Hx2, Hy2, det_H;
    cplx Zxx, Zxy, Zyx, Zyy;
    double cdx = 0, cdy = 0, cdz = 0;
};

// Station data of the device impedance and adjoint right-hand-side kernels
// (m3d_station_desc in include/m3d_device.h).  Per station: whether it has a
// cell, the station cell's edges in cell-edge order with their group role
// (bits 0-1 direction; bit 2 top for x / y edges, +x side for z edges; bit 3
// +y side for z edges: the grouping of :1534-1562) and the cell sizes.  Per
// station edge (sorted, unique): the contributions the adjoint right-hand
// side adds to it, in the order adjoint_rhs adds them (stations in order, in
// each station the E part before the H part, for z edges the y part before
// the x part).  contrib holds 4 ints per contribution: station, kind (0 E_x
// part s_x * we, 1 E_y part s_y * we, 2 H part aH_y * wh, 3 H part aH_x * wh),
// axis * 2 + (sign < 0) with wh = sign * i_axis / n and i_axis =
// 1 / (i omega mu0 h_axis) for axis 0 = z, 1 = x, 2 = y, and n.
struct StationTable {
    std::vector<int> valid;             // [ns]
    std::vector<int> edge_ptr;          // [ns + 1]
    std::vector<int> edge;              // caller edge ids
    std::vector<unsigned char> role;    // per edge of edge
    std::vector<double> cell_size;      // [3 ns] dx, dy, dz
    std::vector<int> uedge;             // [nu] station edges, sorted
    std::vector<int> contrib_ptr;       // [nu + 1]
    std::vector<int> contrib;           // [4 * contributions]
    std::vector<double> contrib_we;     // weight of the E kinds (0 for H kinds)
};

class FVProblem {
public:
    void setup(const OctreeMesh& mesh, const ForwardParams& params);
    void set_sigma(const std::vector<double>& sigma);

    const OctreeMesh& mesh() const { return *mesh_; }
    const StaggeredGrid& grid() const { return grid_; }
    const DiscreteOperators& ops() const { return ops_; }
    const ForwardParams& params() const { return params_; }
    ForwardParams& params() { return params_; }
    int num_edges() const { return ops_.num_edges(); }
    int num_nodes() const { return ops_.num_nodes(); }
    int num_cells() const { return ops_.num_cells(); }

    // Geometry-only flags.
    std::vector<char> is_bc;           // solve Dirichlet set, tol cell_size(0)*0.1
    std::vector<char> is_dead;         // hanging coarse edges
    std::vector<char> is_constrained;  // is_bc | is_dead: identity rows
    std::vector<char> dc_bc_edge;      // DivCorr edge set, tol 1 m, dead excluded
    std::vector<char> dc_bc_node;      // endpoints of dc_bc_edge
    std::vector<char> sens_constrained;  // dead | tol-1 m boundary | air

    // Frequency-independent device operator in split form: off-diagonal
    // stiffness restricted to free rows and free columns, the stiffness
    // diagonal, and identity rows for constrained edges.
    CsrReal offdiag;
    std::vector<double> diag_s;

    // sigma-dependent data.
    std::vector<double> sigma, me, me_s;
    CsrReal L_bc;                 // DivCorr Laplacian after the Dirichlet rewrite
    std::vector<double> L_diag;   // Jacobi diagonal taken before the rewrite

    // Primary-field wavenumber alpha = sqrt(i omega mu0 sigma_bg).
    cplx alpha_bg(double omega) const;
    cplx primary(double omega, double z) const;
    // E_pri on edges of direction `dir` (0 = x for pol 1, 1 = y for pol 2).
    void primary_field(double omega, int dir, std::vector<cplx>& e) const;
    // Dirichlet values for one polarisation (first air-BC pass).
    void bc_values(double omega, int dir, std::vector<cplx>& ebc) const;
    // rhs = -sum_{j in BC} A_ij Ebc_j on free rows, Ebc on BC rows, 0 on dead rows.
    void rhs_total(const std::vector<cplx>& ebc, std::vector<cplx>& rhs) const;
    // phi0 = G^T rhs / (i omega) over non-dead edges.
    void phi0(double omega, const std::vector<cplx>& rhs, std::vector<cplx>& out) const;

    // Station handling.
    void find_stations(const std::vector<std::array<double, 2>>& xy);
    const std::vector<int>& station_cells() const { return station_cell_; }
    void extract_impedance(double omega, const std::vector<cplx>& e1, const std::vector<cplx>& e2,
                           std::vector<std::array<cplx, 4>>& z,
                           std::vector<StationHCache>& cache) const;
    void adjoint_rhs(double omega, const std::vector<StationHCache>& cache,
                     const std::vector<std::array<cplx, 4>>& wr, std::vector<cplx>& rhs1,
                     std::vector<cplx>& rhs2) const;
    // The same two steps on compact arrays over the station cells' edges
    // (station_edges(): sorted, unique), for fields that stay on the device:
    // value k belongs to edge edges[k].  Results equal the full-vector forms.
    std::vector<int> station_edges() const;
    void extract_impedance_compact(double omega, const std::vector<int>& edges, const cplx* e1,
                                   const cplx* e2, std::vector<std::array<cplx, 4>>& z,
                                   std::vector<StationHCache>& cache) const;
    void adjoint_rhs_compact(double omega, const std::vector<StationHCache>& cache,
                             const std::vector<std::array<cplx, 4>>& wr, const std::vector<int>& edges,
                             cplx* rhs1, cplx* rhs2) const;
    void station_table(StationTable& t) const;
    // Conductivity of each station's cell (the 1-D fallback of the
    // impedance, :1704-1712), 0 for stations without a cell.
    void station_sigma(std::vector<double>& s) const;
    // Per-edge weight A_dual * L / n_adj of the sensitivity sum (0 for the
    // edges it skips), the data the device's sensitivity kernel needs.
    void sensitivity_weights(std::vector<double>& w) const;
    void sensitivity(double omega, const std::vector<cplx>& e_bg, const std::vector<cplx>& lambda,
                     std::vector<double>& sens) const;
    void perturbation_rhs(double omega, const std::vector<cplx>& e_bg,
                          const std::vector<double>& delta_sigma, std::vector<cplx>& rhs) const;
    void delta_impedance(const std::vector<cplx>& de, int polarization,
                         const std::vector<StationHCache>& cache,
                         std::vector<std::array<cplx, 4>>& dz) const;
    void delta_impedance_compact(const std::vector<int>& edges, const cplx* de, int polarization,
                                 const std::vector<StationHCache>& cache,
                                 std::vector<std::array<cplx, 4>>& dz) const;

private:
    template <typename Get1, typename Get2>
    void extract_impedance_core(double omega, Get1 e1, Get2 e2, std::vector<std::array<cplx, 4>>& z,
                                std::vector<StationHCache>& cache) const;
    template <typename Get>
    void delta_impedance_core(Get de, int polarization, const std::vector<StationHCache>& cache,
                              std::vector<std::array<cplx, 4>>& dz) const;
    template <typename Add>
    void adjoint_rhs_core(double omega, const std::vector<StationHCache>& cache,
                          const std::vector<std::array<cplx, 4>>& wr, Add add) const;

    const OctreeMesh* mesh_ = nullptr;
    ForwardParams params_;
    StaggeredGrid grid_;
    DiscreteOperators ops_;
    double sigma_bg_ = 0.01;
    std::vector<int> station_cell_;
};

}  // namespace m3d
