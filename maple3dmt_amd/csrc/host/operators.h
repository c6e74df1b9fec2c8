// Discrete operators on the octree staggered grid and the frequency-independent
// parts of the curl-curl system A = Ce^T diag(Mf^-1) Ce + i*omega*diag(Me).
//
// Follows DiscreteOperators (src/octree/operators.cpp): geometric caches
// (:19-170), geometric curl with +-L_edge entries (:176-286), topological
// divergence (:292-330), SimPEG-style hanging-face rewrite (:344-466), edge
// mass (:481-518), face mass inverse (:539-551), stiffness assembly in face
// order with a sorted per-row pattern (:556-837), node gradient G_node
// (:849-944), DivCorr Laplacian G^T Me G (:952-1019), div(sigma E) (:1025-1057)
// and G*phi (:1063-1074).
//
// The stiffness S is kept real and separate from the diagonal mass so that a
// device operator can be built once per mesh and shared by every frequency.
#pragma once

#include <complex>
#include <vector>

#include "staggered_grid.h"

namespace m3d {

struct CsrReal {
    int nrows = 0, ncols = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
    int nnz() const { return static_cast<int>(col.size()); }
};

struct HangingPair {
    int coarse, child1, child2;
};

class DiscreteOperators {
public:
    void build(const StaggeredGrid& grid);

    int num_edges() const { return ne_; }
    int num_faces() const { return nf_; }
    int num_cells() const { return nc_; }
    int num_nodes() const { return nn_; }

    const CsrReal& curl() const { return ce_; }          // faces x edges
    const CsrReal& divergence() const { return d_; }     // cells x faces
    const CsrReal& gradient_node() const { return g_; }  // edges x nodes, 2 per row
    // S = Ce^T diag(Mf^-1) Ce with a diagonal entry in every row, columns
    // sorted per row; values accumulated in face order.
    const CsrReal& stiffness() const { return s_; }

    bool is_dead_edge(int e) const { return dead_edge_[e] != 0; }
    bool is_dead_face(int f) const { return dead_face_[f] != 0; }
    const std::vector<HangingPair>& hanging_pairs() const { return hanging_pairs_; }
    double edge_dual_area(int e) const { return edge_dual_area_[e]; }
    const std::vector<double>& edge_dual_areas() const { return edge_dual_area_; }
    const std::vector<double>& edge_lengths() const { return edge_len_; }
    // Cells adjacent to each edge, in the order the edge mass averages them.
    const std::vector<std::vector<int>>& edge_adjacency() const { return edge_adj_; }
    double face_dual_length(int f) const { return face_dual_len_[f]; }
    const std::vector<double>& node_x() const { return node_x_; }
    const std::vector<double>& node_y() const { return node_y_; }
    const std::vector<double>& node_z() const { return node_z_; }
    const std::vector<double>& cell_cx() const { return cell_cx_; }
    const std::vector<double>& cell_cy() const { return cell_cy_; }
    const std::vector<double>& cell_cz() const { return cell_cz_; }

    // Me[e] = mean(sigma_adj) * (A_dual * L), 0 on dead edges (:481-518).
    void edge_mass(const std::vector<double>& sigma, std::vector<double>& me) const;
    // Me_sigma[e] = (mean(sigma_adj) * A_dual) * L, 0 on dead edges; the
    // weight used by the DivCorr Laplacian and div(sigma E) (:956-973, :1034-1048).
    void edge_mass_divcorr(const std::vector<double>& sigma, std::vector<double>& me_s) const;
    // Mf^-1[f] = 1 / (mu0 A_f / L_dual), 0 on dead faces (:539-551).
    void face_mass_inv(std::vector<double>& mfi) const;
    // L = G^T diag(me_sigma) G on nodes, sorted CSR (:952-1019).
    void div_laplacian(const std::vector<double>& me_sigma, CsrReal& L) const;

private:
    int ne_ = 0, nf_ = 0, nc_ = 0, nn_ = 0;
    std::vector<double> edge_len_, edge_x_, edge_y_, edge_z_;
    std::vector<int> edge_dir_;
    std::vector<std::vector<int>> edge_adj_;
    std::vector<double> face_area_;
    std::vector<int> face_normal_, face_minus_, face_plus_;
    std::vector<double> cell_cx_, cell_cy_, cell_cz_, cell_hx_, cell_hy_, cell_hz_;
    std::vector<double> face_dual_len_, edge_dual_area_;
    std::vector<char> dead_edge_, dead_face_;
    std::vector<HangingPair> hanging_pairs_;
    std::vector<double> node_x_, node_y_, node_z_;
    CsrReal ce_, d_, g_, s_;

    void build_curl(const StaggeredGrid& grid);
    void build_divergence(const StaggeredGrid& grid);
    void handle_hanging(const StaggeredGrid& grid);
    void build_node_gradient();
    void build_stiffness();
};

}  // namespace m3d
