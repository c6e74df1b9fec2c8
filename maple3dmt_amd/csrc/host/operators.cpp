// Discrete operators; see operators.h for the reference mapping.
#include "operators.h"
#include "flat_map.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>

namespace m3d {

namespace {

// Build a CSR from (row, col, val) triplets keeping insertion order per row.
CsrReal triplets_to_csr(int nrows, int ncols, const std::vector<int>& r,
                        const std::vector<int>& c, const std::vector<double>& v) {
    CsrReal m;
    m.nrows = nrows;
    m.ncols = ncols;
    m.ptr.assign(nrows + 1, 0);
    for (int row : r) m.ptr[row + 1]++;
    for (int i = 0; i < nrows; ++i) m.ptr[i + 1] += m.ptr[i];
    m.col.resize(r.size());
    m.val.resize(r.size());
    std::vector<int> pos(m.ptr.begin(), m.ptr.end() - 1);
    for (size_t k = 0; k < r.size(); ++k) {
        const int p = pos[r[k]]++;
        m.col[p] = c[k];
        m.val[p] = v[k];
    }
    return m;
}

// Sorted, de-duplicated pattern from per-row column lists.
void pattern_from_rows(std::vector<std::vector<int>>& rows, CsrReal& m) {
    m.nrows = static_cast<int>(rows.size());
    m.ptr.assign(m.nrows + 1, 0);
    for (int i = 0; i < m.nrows; ++i) {
        auto& r = rows[i];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        m.ptr[i + 1] = m.ptr[i] + static_cast<int>(r.size());
    }
    m.col.resize(m.ptr[m.nrows]);
    for (int i = 0; i < m.nrows; ++i) std::copy(rows[i].begin(), rows[i].end(), m.col.begin() + m.ptr[i]);
    m.val.assign(m.col.size(), 0.0);
}

inline int find_col(const CsrReal& m, int row, int col) {
    auto b = m.col.begin() + m.ptr[row], e = m.col.begin() + m.ptr[row + 1];
    auto it = std::lower_bound(b, e, col);
    return (it != e && *it == col) ? static_cast<int>(it - m.col.begin()) : -1;
}

}  // namespace

void DiscreteOperators::build(const StaggeredGrid& grid) {
    ne_ = grid.num_edges();
    nf_ = grid.num_faces();
    nc_ = grid.num_cells();
    edge_len_ = grid.edge_len;
    edge_x_ = grid.edge_x;
    edge_y_ = grid.edge_y;
    edge_z_ = grid.edge_z;
    edge_dir_.assign(grid.edge_dir.begin(), grid.edge_dir.end());
    edge_adj_ = grid.edge_adj;
    face_area_ = grid.face_area;
    face_normal_.assign(grid.face_normal.begin(), grid.face_normal.end());
    face_minus_ = grid.face_minus;
    face_plus_ = grid.face_plus;

    // Cell centres as the mean of adjacent edge midpoints, accumulated in edge
    // order (:60-86).  Coarse cells at hanging faces also count the fine edges
    // added to their adjacency, exactly as the reference does.
    cell_cx_.assign(nc_, 0.0); cell_cy_.assign(nc_, 0.0); cell_cz_.assign(nc_, 0.0);
    cell_hx_.assign(nc_, 0.0); cell_hy_.assign(nc_, 0.0); cell_hz_.assign(nc_, 0.0);
    std::vector<int> cnt(nc_, 0);
    for (int e = 0; e < ne_; ++e) {
        for (int c : edge_adj_[e]) {
            if (c < 0 || c >= nc_) continue;
            cell_cx_[c] += edge_x_[e];
            cell_cy_[c] += edge_y_[e];
            cell_cz_[c] += edge_z_[e];
            cnt[c]++;
        }
    }
    for (int c = 0; c < nc_; ++c) {
        if (cnt[c] > 0) {
            cell_cx_[c] /= cnt[c];
            cell_cy_[c] /= cnt[c];
            cell_cz_[c] /= cnt[c];
        }
    }
    // Half-extents: largest half edge length per direction (:90-99).
    for (int e = 0; e < ne_; ++e) {
        const double half = edge_len_[e] / 2.0;
        const int d = edge_dir_[e];
        for (int c : edge_adj_[e]) {
            if (c < 0 || c >= nc_) continue;
            double& h = (d == 0) ? cell_hx_[c] : (d == 1) ? cell_hy_[c] : cell_hz_[c];
            if (half > h) h = half;
        }
    }
    // Face dual lengths (:103-126).
    face_dual_len_.resize(nf_);
    for (int f = 0; f < nf_; ++f) {
        const int n = face_normal_[f], cm = face_minus_[f], cp = face_plus_[f];
        auto coord = [&](int c) { return n == 0 ? cell_cx_[c] : n == 1 ? cell_cy_[c] : cell_cz_[c]; };
        auto half = [&](int c) { return n == 0 ? cell_hx_[c] : n == 1 ? cell_hy_[c] : cell_hz_[c]; };
        double L;
        if (cm >= 0 && cp >= 0) {
            L = std::abs(coord(cp) - coord(cm));
        } else {
            const int c = (cm >= 0) ? cm : cp;
            L = (c >= 0) ? half(c) : std::sqrt(face_area_[f]);
        }
        if (L < 1e-20) L = std::sqrt(face_area_[f]);
        face_dual_len_[f] = L;
    }
    // Edge dual areas: sum of per-cell perpendicular half-extent products
    // (:133-157).
    edge_dual_area_.resize(ne_);
    for (int e = 0; e < ne_; ++e) {
        const auto& adj = edge_adj_[e];
        const int d = edge_dir_[e];
        double sum = 0;
        int count = 0;
        for (int c : adj) {
            if (c < 0 || c >= nc_) continue;
            const double p1 = (d == 0) ? cell_hy_[c] : cell_hx_[c];
            const double p2 = (d == 2) ? cell_hy_[c] : cell_hz_[c];
            sum += p1 * p2;
            ++count;
        }
        edge_dual_area_[e] = (count > 0) ? sum : edge_len_[e] * edge_len_[e];
    }

    build_curl(grid);
    build_divergence(grid);
    handle_hanging(grid);
    build_node_gradient();
    build_stiffness();
}

// Geometric curl: Ce[f,e] = sign * L_e with the right-hand circulation signs
// of :208-235.
void DiscreteOperators::build_curl(const StaggeredGrid& grid) {
    std::vector<int> r, c;
    std::vector<double> v;
    r.reserve(size_t(nf_) * 4); c.reserve(size_t(nf_) * 4); v.reserve(size_t(nf_) * 4);
    for (int f = 0; f < nf_; ++f) {
        const int n = grid.face_normal[f];
        for (int e : grid.face_edges[f]) {
            if (e < 0 || e >= ne_) continue;
            const int d = grid.edge_dir[e];
            double sign = 0.0;
            if (n == 0) {
                if (d == 1) sign = (grid.edge_z[e] < grid.face_z[f]) ? 1.0 : -1.0;
                else if (d == 2) sign = (grid.edge_y[e] < grid.face_y[f]) ? -1.0 : 1.0;
            } else if (n == 1) {
                if (d == 2) sign = (grid.edge_x[e] < grid.face_x[f]) ? 1.0 : -1.0;
                else if (d == 0) sign = (grid.edge_z[e] < grid.face_z[f]) ? -1.0 : 1.0;
            } else {
                if (d == 0) sign = (grid.edge_y[e] < grid.face_y[f]) ? 1.0 : -1.0;
                else if (d == 1) sign = (grid.edge_x[e] < grid.face_x[f]) ? -1.0 : 1.0;
            }
            if (std::abs(sign) > 0.5) {
                r.push_back(f);
                c.push_back(e);
                v.push_back(sign * edge_len_[e]);
            }
        }
    }
    ce_ = triplets_to_csr(nf_, ne_, r, c, v);
}

// Topological divergence D[c,f] = +1 (minus side) / -1 (plus side), :292-330.
void DiscreteOperators::build_divergence(const StaggeredGrid& grid) {
    std::vector<int> r, c;
    std::vector<double> v;
    for (int f = 0; f < nf_; ++f) {
        if (grid.face_minus[f] >= 0 && grid.face_minus[f] < nc_) {
            r.push_back(grid.face_minus[f]); c.push_back(f); v.push_back(1.0);
        }
        if (grid.face_plus[f] >= 0 && grid.face_plus[f] < nc_) {
            r.push_back(grid.face_plus[f]); c.push_back(f); v.push_back(-1.0);
        }
    }
    d_ = triplets_to_csr(nc_, nf_, r, c, v);
}

// Dead coarse faces lose their Ce rows; dead coarse edges are replaced by
// 0.5*(child1 + child2) in alive rows, merged per (row, col) and re-sorted
// (:344-466).
void DiscreteOperators::handle_hanging(const StaggeredGrid& grid) {
    dead_edge_.assign(ne_, 0);
    dead_face_.assign(nf_, 0);
    hanging_pairs_.clear();
    if (grid.hanging_faces.empty()) return;

    std::unordered_map<int, std::pair<int, int>> children;
    for (const auto& he : grid.hanging_edges) {
        dead_edge_[he.coarse_edge] = 1;
        children[he.coarse_edge] = {he.fine_edges[0], he.fine_edges[1]};
        hanging_pairs_.push_back({he.coarse_edge, he.fine_edges[0], he.fine_edges[1]});
    }
    for (const auto& hf : grid.hanging_faces) dead_face_[hf.coarse_face] = 1;

    CsrReal out;
    out.nrows = nf_;
    out.ncols = ne_;
    out.ptr.assign(nf_ + 1, 0);
    std::vector<std::pair<int, double>> row;
    for (int f = 0; f < nf_; ++f) {
        row.clear();
        if (!dead_face_[f]) {
            for (int k = ce_.ptr[f]; k < ce_.ptr[f + 1]; ++k) {
                const int e = ce_.col[k];
                const double val = ce_.val[k];
                if (std::abs(val) < 1e-30) continue;
                if (dead_edge_[e]) {
                    const auto ch = children[e];
                    row.push_back({ch.first, 0.5 * val});
                    row.push_back({ch.second, 0.5 * val});
                } else {
                    row.push_back({e, val});
                }
            }
        }
        // Merge equal columns (sum in insertion order) and sort by column.
        std::map<int, double> merged;
        for (const auto& kv : row) merged[kv.first] += kv.second;
        for (const auto& kv : merged) {
            if (std::abs(kv.second) > 1e-30) {
                out.col.push_back(kv.first);
                out.val.push_back(kv.second);
            }
        }
        out.ptr[f + 1] = static_cast<int>(out.col.size());
    }
    ce_ = std::move(out);
}

// Nodes from edge endpoints quantised at 1% of the shortest edge, numbered
// by first appearance; G[e,v1] = -1/L, G[e,v2] = +1/L stored in column order
// (:849-944).
void DiscreteOperators::build_node_gradient() {
    double min_L = 1e30;
    for (int e = 0; e < ne_; ++e)
        if (edge_len_[e] > 1e-10) min_L = std::min(min_L, edge_len_[e]);
    const double tol = min_L * 0.01;
    auto q = [&](double v) { return static_cast<int64_t>(std::round(v / tol)); };

    struct Key {
        int64_t x, y, z;
        bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash {
        uint64_t operator()(const Key& k) const {
            return mix64(uint64_t(k.x) ^ mix64(uint64_t(k.y) ^ mix64(uint64_t(k.z))));
        }
    };
    FlatMap<Key, KeyHash> nodes(size_t(ne_) / 2 + 16);
    std::vector<Key> keys;
    auto node_of = [&](double x, double y, double z) {
        const Key k{q(x), q(y), q(z)};
        const auto r = nodes.insert(k, static_cast<int>(keys.size()));
        if (r.second) keys.push_back(k);
        return r.first;
    };
    std::vector<int> n1(ne_), n2(ne_);
    for (int e = 0; e < ne_; ++e) {
        const double hl = edge_len_[e] * 0.5;
        double a[3] = {edge_x_[e], edge_y_[e], edge_z_[e]};
        double b[3] = {edge_x_[e], edge_y_[e], edge_z_[e]};
        a[edge_dir_[e]] -= hl;
        b[edge_dir_[e]] += hl;
        n1[e] = node_of(a[0], a[1], a[2]);
        n2[e] = node_of(b[0], b[1], b[2]);
    }
    nn_ = static_cast<int>(keys.size());
    node_x_.resize(nn_); node_y_.resize(nn_); node_z_.resize(nn_);
    for (int i = 0; i < nn_; ++i) {
        node_x_[i] = keys[i].x * tol;
        node_y_[i] = keys[i].y * tol;
        node_z_[i] = keys[i].z * tol;
    }
    g_.nrows = ne_;
    g_.ncols = nn_;
    g_.ptr.resize(ne_ + 1);
    for (int e = 0; e <= ne_; ++e) g_.ptr[e] = 2 * e;
    g_.col.resize(2 * size_t(ne_));
    g_.val.resize(2 * size_t(ne_));
    for (int e = 0; e < ne_; ++e) {
        const double inv_L = 1.0 / edge_len_[e];
        if (n1[e] < n2[e]) {
            g_.col[2 * e] = n1[e]; g_.val[2 * e] = -inv_L;
            g_.col[2 * e + 1] = n2[e]; g_.val[2 * e + 1] = +inv_L;
        } else {
            g_.col[2 * e] = n2[e]; g_.val[2 * e] = +inv_L;
            g_.col[2 * e + 1] = n1[e]; g_.val[2 * e + 1] = -inv_L;
        }
    }
}

void DiscreteOperators::face_mass_inv(std::vector<double>& mfi) const {
    mfi.resize(nf_);
    for (int f = 0; f < nf_; ++f) {
        if (dead_face_[f]) { mfi[f] = 0.0; continue; }
        const double mf = kMu0 * face_area_[f] / face_dual_len_[f];
        mfi[f] = (mf > 1e-30) ? 1.0 / mf : 0.0;
    }
}

// S = Ce^T diag(Mf^-1) Ce.  Pattern: every (i, j) sharing a face row of Ce
// plus the diagonal; values summed over faces in face order as
// (Ce[f,i] * Mf^-1[f]) * Ce[f,j] (:586-604).
void DiscreteOperators::build_stiffness() {
    std::vector<double> mfi;
    face_mass_inv(mfi);
    std::vector<std::vector<int>> rows(ne_);
    for (int e = 0; e < ne_; ++e) rows[e].push_back(e);
    for (int f = 0; f < nf_; ++f) {
        for (int k1 = ce_.ptr[f]; k1 < ce_.ptr[f + 1]; ++k1)
            for (int k2 = ce_.ptr[f]; k2 < ce_.ptr[f + 1]; ++k2)
                rows[ce_.col[k1]].push_back(ce_.col[k2]);
    }
    pattern_from_rows(rows, s_);
    s_.ncols = ne_;
    for (int f = 0; f < nf_; ++f) {
        const double m = mfi[f];
        for (int k1 = ce_.ptr[f]; k1 < ce_.ptr[f + 1]; ++k1) {
            const int i = ce_.col[k1];
            const double a = ce_.val[k1];
            for (int k2 = ce_.ptr[f]; k2 < ce_.ptr[f + 1]; ++k2) {
                const int idx = find_col(s_, i, ce_.col[k2]);
                s_.val[idx] += a * m * ce_.val[k2];
            }
        }
    }
}

void DiscreteOperators::edge_mass(const std::vector<double>& sigma, std::vector<double>& me) const {
    me.assign(ne_, 0.0);
    const int ns = static_cast<int>(sigma.size());
    for (int e = 0; e < ne_; ++e) {
        const auto& adj = edge_adj_[e];
        if (adj.empty()) continue;
        double sum = 0.0;
        int count = 0;
        for (int c : adj)
            if (c >= 0 && c < ns) { sum += sigma[c]; ++count; }
        const double avg = (count > 0) ? sum / count : 1e-8;
        const double hodge = edge_dual_area_[e] * edge_len_[e];
        me[e] = avg * hodge;
    }
    for (const auto& hp : hanging_pairs_) me[hp.coarse] = 0.0;
}

void DiscreteOperators::edge_mass_divcorr(const std::vector<double>& sigma,
                                          std::vector<double>& me_s) const {
    me_s.assign(ne_, 0.0);
    const int ns = static_cast<int>(sigma.size());
    for (int e = 0; e < ne_; ++e) {
        if (dead_edge_[e]) continue;
        double sum = 0.0;
        int count = 0;
        for (int c : edge_adj_[e])
            if (c >= 0 && c < ns) { sum += sigma[c]; ++count; }
        const double avg = (count > 0) ? sum / count : 1e-8;
        me_s[e] = avg * edge_dual_area_[e] * edge_len_[e];
    }
}

void DiscreteOperators::div_laplacian(const std::vector<double>& me_sigma, CsrReal& L) const {
    std::vector<std::vector<int>> rows(nn_);
    for (int e = 0; e < ne_; ++e) {
        if (me_sigma[e] < 1e-30) continue;
        const int a = g_.col[2 * e], b = g_.col[2 * e + 1];
        rows[a].push_back(a); rows[a].push_back(b);
        rows[b].push_back(a); rows[b].push_back(b);
    }
    pattern_from_rows(rows, L);
    L.ncols = nn_;
    for (int e = 0; e < ne_; ++e) {
        const double w = me_sigma[e];
        if (w < 1e-30) continue;
        for (int k1 = 2 * e; k1 < 2 * e + 2; ++k1) {
            const int r = g_.col[k1];
            const double g1 = g_.val[k1];
            for (int k2 = 2 * e; k2 < 2 * e + 2; ++k2) {
                const int idx = find_col(L, r, g_.col[k2]);
                L.val[idx] += g1 * w * g_.val[k2];
            }
        }
    }
}

}  // namespace m3d
