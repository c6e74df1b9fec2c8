// Host-side parts of the FV solve; see fv_problem.h for the reference mapping.
#include "fv_problem.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace m3d {

void FVProblem::setup(const OctreeMesh& mesh, const ForwardParams& params) {
    mesh_ = &mesh;
    params_ = params;
    sigma_bg_ = mesh.params().sigma_bg;
    if (sigma_bg_ <= 0) sigma_bg_ = 0.01;  // forward_solver_fv.cpp:34-35
    grid_.build(mesh);
    ops_.build(grid_);

    const int ne = ops_.num_edges();
    const auto& mp = mesh.params();
    const double air_z = params_.air_z_threshold;

    // Solve Dirichlet set: six domain faces within cell_size(0)*0.1 plus every
    // air edge above air_z_threshold (forward_solver_fv.cpp:117, 166-194).
    const double btol = mesh.cell_size(0) * 0.1;
    is_bc.assign(ne, 0);
    is_dead.assign(ne, 0);
    is_constrained.assign(ne, 0);
    for (int e = 0; e < ne; ++e) {
        const double x = grid_.edge_x[e], y = grid_.edge_y[e], z = grid_.edge_z[e];
        const bool on_boundary = x <= mp.domain_x_min + btol || x >= mp.domain_x_max - btol ||
                                 y <= mp.domain_y_min + btol || y >= mp.domain_y_max - btol ||
                                 z <= mp.domain_z_min + btol || z >= mp.domain_z_max - btol;
        if (on_boundary || z > air_z) is_bc[e] = 1;
    }
    for (const auto& hp : ops_.hanging_pairs()) is_dead[hp.coarse] = 1;
    for (int e = 0; e < ne; ++e) is_constrained[e] = is_bc[e] || is_dead[e];

    // DivCorr and sensitivity sets use a 1 m tolerance (:556-569, :1017-1029).
    dc_bc_edge.assign(ne, 0);
    sens_constrained.assign(ne, 0);
    for (int e = 0; e < ne; ++e) {
        const double x = grid_.edge_x[e], y = grid_.edge_y[e], z = grid_.edge_z[e];
        const double t = 1.0;
        const bool outer = x <= mp.domain_x_min + t || x >= mp.domain_x_max - t ||
                           y <= mp.domain_y_min + t || y >= mp.domain_y_max - t ||
                           z <= mp.domain_z_min + t || z >= mp.domain_z_max - t || z > air_z;
        if (ops_.is_dead_edge(e)) {
            sens_constrained[e] = 1;
            continue;
        }
        if (outer) {
            dc_bc_edge[e] = 1;
            sens_constrained[e] = 1;
        }
    }
    const CsrReal& G = ops_.gradient_node();
    dc_bc_node.assign(ops_.num_nodes(), 0);
    for (int e = 0; e < ne; ++e)
        if (dc_bc_edge[e])
            for (int k = G.ptr[e]; k < G.ptr[e + 1]; ++k) dc_bc_node[G.col[k]] = 1;

    // Split device operator: free rows keep their couplings to free columns;
    // constrained rows become identity and constrained columns drop out
    // (the identity elimination of :224-235).
    const CsrReal& S = ops_.stiffness();
    offdiag.nrows = offdiag.ncols = ne;
    offdiag.ptr.assign(ne + 1, 0);
    offdiag.col.clear();
    offdiag.val.clear();
    diag_s.assign(ne, 1.0);
    for (int i = 0; i < ne; ++i) {
        if (!is_constrained[i]) {
            for (int k = S.ptr[i]; k < S.ptr[i + 1]; ++k) {
                const int j = S.col[k];
                if (j == i) {
                    diag_s[i] = S.val[k];
                } else if (!is_constrained[j]) {
                    offdiag.col.push_back(j);
                    offdiag.val.push_back(S.val[k]);
                }
            }
        }
        offdiag.ptr[i + 1] = static_cast<int>(offdiag.col.size());
    }
}

void FVProblem::set_sigma(const std::vector<double>& s) {
    if (static_cast<int>(s.size()) != ops_.num_cells())
        throw std::invalid_argument("sigma size does not match the number of cells");
    sigma = s;
    ops_.edge_mass(sigma, me);
    ops_.edge_mass_divcorr(sigma, me_s);

    // DivCorr Laplacian with phi = 0 Dirichlet rows and columns (:581-606).
    ops_.div_laplacian(me_s, L_bc);
    const int nn = ops_.num_nodes();
    L_diag.assign(nn, 1.0);
    for (int i = 0; i < nn; ++i) {
        if (dc_bc_node[i]) continue;
        for (int k = L_bc.ptr[i]; k < L_bc.ptr[i + 1]; ++k)
            if (L_bc.col[k] == i) { L_diag[i] = L_bc.val[k]; break; }
    }
    for (int i = 0; i < nn; ++i) {
        if (!dc_bc_node[i]) continue;
        for (int k = L_bc.ptr[i]; k < L_bc.ptr[i + 1]; ++k) L_bc.val[k] = (L_bc.col[k] == i) ? 1.0 : 0.0;
    }
    for (int i = 0; i < nn; ++i) {
        if (dc_bc_node[i]) continue;
        for (int k = L_bc.ptr[i]; k < L_bc.ptr[i + 1]; ++k)
            if (dc_bc_node[L_bc.col[k]]) L_bc.val[k] = 0.0;
    }
}

cplx FVProblem::alpha_bg(double omega) const {
    return std::sqrt(cplx(0.0, omega * kMu0 * sigma_bg_));
}

// E = 1 + alpha z in air, exp(alpha z) in the earth (:133-136).
cplx FVProblem::primary(double omega, double z) const {
    const cplx a = alpha_bg(omega);
    if (z >= 0.0) return cplx(1.0, 0.0) + a * z;
    return std::exp(a * z);
}

void FVProblem::primary_field(double omega, int dir, std::vector<cplx>& e) const {
    const int ne = num_edges();
    e.assign(ne, cplx(0, 0));
    const cplx a = alpha_bg(omega);
    for (int i = 0; i < ne; ++i) {
        if (grid_.edge_dir[i] != dir) continue;
        const double z = grid_.edge_z[i];
        e[i] = (z >= 0.0) ? cplx(1.0, 0.0) + a * z : std::exp(a * z);
    }
}

void FVProblem::bc_values(double omega, int dir, std::vector<cplx>& ebc) const {
    const int ne = num_edges();
    ebc.assign(ne, cplx(0, 0));
    for (int i = 0; i < ne; ++i)
        if (is_bc[i] && grid_.edge_dir[i] == dir) ebc[i] = primary(omega, grid_.edge_z[i]);
}

void FVProblem::rhs_total(const std::vector<cplx>& ebc, std::vector<cplx>& rhs) const {
    const int ne = num_edges();
    const CsrReal& S = ops_.stiffness();
    rhs.assign(ne, cplx(0, 0));
    for (int i = 0; i < ne; ++i) {
        if (is_constrained[i]) continue;
        for (int k = S.ptr[i]; k < S.ptr[i + 1]; ++k) {
            const int j = S.col[k];
            if (j != i && is_bc[j]) rhs[i] -= cplx(S.val[k], 0.0) * ebc[j];
        }
    }
    for (int i = 0; i < ne; ++i)
        if (is_bc[i]) rhs[i] = ebc[i];
}

void FVProblem::phi0(double omega, const std::vector<cplx>& rhs, std::vector<cplx>& out) const {
    const CsrReal& G = ops_.gradient_node();
    out.assign(ops_.num_nodes(), cplx(0, 0));
    for (int e = 0; e < num_edges(); ++e) {
        if (ops_.is_dead_edge(e)) continue;
        for (int k = G.ptr[e]; k < G.ptr[e + 1]; ++k) out[G.col[k]] += G.val[k] * rhs[e];
    }
    const cplx iw(0.0, omega);
    for (auto& v : out) v /= iw;
}

// Shallowest non-air cell whose horizontal extent (+-1 m) covers the station,
// first in cell order; else the nearest centre (:1393-1485).
void FVProblem::find_stations(const std::vector<std::array<double, 2>>& xy) {
    station_cell_.clear();
    const OctreeMesh& m = *mesh_;
    for (const auto& st : xy) {
        const double sx = st[0], sy = st[1], sz = 0.0;
        int best = -1, fallback = -1;
        double best_cz = 1e30, fb_dist = 1e30;
        for (int c = 0; c < m.num_cells(); ++c) {
            if (m.cell_type(c) == CellType::AIR) continue;
            double cx, cy, cz, hx, hy, hz;
            m.cell_center(c, cx, cy, cz);
            m.cell_size_xyz(c, hx, hy, hz);
            const bool in_x = (sx >= cx - hx / 2 - 1.0) && (sx <= cx + hx / 2 + 1.0);
            const bool in_y = (sy >= cy - hy / 2 - 1.0) && (sy <= cy + hy / 2 + 1.0);
            if (in_x && in_y && cz < 0 && std::abs(cz) < best_cz) {
                best_cz = std::abs(cz);
                best = c;
            }
            const double d = std::sqrt((cx - sx) * (cx - sx) + (cy - sy) * (cy - sy) + (cz - sz) * (cz - sz));
            if (d < fb_dist) {
                fb_dist = d;
                fallback = c;
            }
        }
        station_cell_.push_back(best >= 0 ? best : fallback);
    }
}

namespace {

// Edge groups of one station cell: x/y edges split top/bottom by z against
// the cell centre, z edges split +-x and +-y (:1534-1562).
struct CellGroups {
    int nxt = 0, nxb = 0, nyt = 0, nyb = 0, nzpx = 0, nzmx = 0, nzpy = 0, nzmy = 0;
};

}  // namespace

template <typename Get1, typename Get2>
void FVProblem::extract_impedance_core(double omega, Get1 e1, Get2 e2, std::vector<std::array<cplx, 4>>& z,
                                       std::vector<StationHCache>& cache) const {
    const OctreeMesh& m = *mesh_;
    const cplx iwmu(0.0, omega * kMu0);
    const int ns = static_cast<int>(station_cell_.size());
    z.assign(ns, {cplx(0, 0), cplx(0, 0), cplx(0, 0), cplx(0, 0)});
    cache.assign(ns, StationHCache{});
    for (int s = 0; s < ns; ++s) {
        const int cell = station_cell_[s];
        if (cell < 0) continue;
        double cx, cy, cz, cdx, cdy, cdz;
        m.cell_center(cell, cx, cy, cz);
        m.cell_size_xyz(cell, cdx, cdy, cdz);

        cplx x1t, x1b, x2t, x2b, y1t, y1b, y2t, y2b;
        cplx z1px, z1mx, z2px, z2mx, z1py, z1my, z2py, z2my;
        CellGroups g;
        for (const CellEdge& ce : grid_.cell_edges[cell]) {
            const int e = ce.edge_id;
            const int d = grid_.edge_dir[e];
            if (d == 0) {
                if (grid_.edge_z[e] > cz) { x1t += e1(e); x2t += e2(e); ++g.nxt; }
                else { x1b += e1(e); x2b += e2(e); ++g.nxb; }
            } else if (d == 1) {
                if (grid_.edge_z[e] > cz) { y1t += e1(e); y2t += e2(e); ++g.nyt; }
                else { y1b += e1(e); y2b += e2(e); ++g.nyb; }
            } else {
                if (grid_.edge_x[e] > cx) { z1px += e1(e); z2px += e2(e); ++g.nzpx; }
                else { z1mx += e1(e); z2mx += e2(e); ++g.nzmx; }
                if (grid_.edge_y[e] > cy) { z1py += e1(e); z2py += e2(e); ++g.nzpy; }
                else { z1my += e1(e); z2my += e2(e); ++g.nzmy; }
            }
        }
        auto avg = [](cplx& a, cplx& b, int n) {
            if (n > 0) { a /= cplx(n, 0); b /= cplx(n, 0); }
        };
        avg(x1t, x2t, g.nxt); avg(x1b, x2b, g.nxb); avg(y1t, y2t, g.nyt); avg(y1b, y2b, g.nyb);
        avg(z1px, z2px, g.nzpx); avg(z1mx, z2mx, g.nzmx); avg(z1py, z2py, g.nzpy); avg(z1my, z2my, g.nzmy);

        auto centre = [](const cplx& t, const cplx& b, int nt, int nb) {
            if (nt > 0 && nb > 0) return (t + b) * cplx(0.5, 0);
            return nt > 0 ? t : b;
        };
        const cplx Ex1 = centre(x1t, x1b, g.nxt, g.nxb);
        const cplx Ey1 = centre(y1t, y1b, g.nyt, g.nyb);
        const cplx Ex2 = centre(x2t, x2b, g.nxt, g.nxb);
        const cplx Ey2 = centre(y2t, y2b, g.nyt, g.nyb);

        cplx dEx1dz, dEx2dz, dEy1dz, dEy2dz, dEz1dx, dEz2dx, dEz1dy, dEz2dy;
        if (g.nxt > 0 && g.nxb > 0) { dEx1dz = (x1t - x1b) / cdz; dEx2dz = (x2t - x2b) / cdz; }
        if (g.nyt > 0 && g.nyb > 0) { dEy1dz = (y1t - y1b) / cdz; dEy2dz = (y2t - y2b) / cdz; }
        if (g.nzpx > 0 && g.nzmx > 0) { dEz1dx = (z1px - z1mx) / cdx; dEz2dx = (z2px - z2mx) / cdx; }
        if (g.nzpy > 0 && g.nzmy > 0) { dEz1dy = (z1py - z1my) / cdy; dEz2dy = (z2py - z2my) / cdy; }

        // Faraday: H = curl(E) / (i omega mu0), :1629-1632.
        const cplx Hx1 = (dEz1dy - dEy1dz) / iwmu;
        const cplx Hy1 = (dEx1dz - dEz1dx) / iwmu;
        const cplx Hx2 = (dEz2dy - dEy2dz) / iwmu;
        const cplx Hy2 = (dEx2dz - dEz2dx) / iwmu;
        const cplx det = Hx1 * Hy2 - Hx2 * Hy1;

        StationHCache& hc = cache[s];
        hc.Hx1 = Hx1; hc.Hy1 = Hy1; hc.Hx2 = Hx2; hc.Hy2 = Hy2; hc.det_H = det;
        hc.cdx = cdx; hc.cdy = cdy; hc.cdz = cdz;
        cplx Zxx, Zxy, Zyx, Zyy;
        if (std::abs(det) > 1e-30) {
            Zxx = (Ex1 * Hy2 - Ex2 * Hy1) / det;
            Zxy = (Ex2 * Hx1 - Ex1 * Hx2) / det;
            Zyx = (Ey1 * Hy2 - Ey2 * Hy1) / det;
            Zyy = (Ey2 * Hx1 - Ey1 * Hx2) / det;
        } else {
            const cplx z1d = std::sqrt(iwmu / cplx(sigma[cell], 0));
            Zxx = cplx(0, 0); Zxy = z1d; Zyx = -z1d; Zyy = cplx(0, 0);
        }
        hc.Zxx = Zxx; hc.Zxy = Zxy; hc.Zyx = Zyx; hc.Zyy = Zyy;
        z[s] = {Zxx, Zxy, Zyx, Zyy};
    }
}

void FVProblem::extract_impedance(double omega, const std::vector<cplx>& e1,
                                  const std::vector<cplx>& e2,
                                  std::vector<std::array<cplx, 4>>& z,
                                  std::vector<StationHCache>& cache) const {
    auto g1 = [&](int e) -> const cplx& { return e1[e]; };
    auto g2 = [&](int e) -> const cplx& { return e2[e]; };
    extract_impedance_core(omega, g1, g2, z, cache);
}

// Q^T of Z = E H^-1: E part via H^-T and H part via -Z^T and the transpose
// of the station-cell curl (:1804-2013).
template <typename Add>
void FVProblem::adjoint_rhs_core(double omega, const std::vector<StationHCache>& cache,
                                 const std::vector<std::array<cplx, 4>>& wr, Add add) const {
    const cplx iwmu(0.0, omega * kMu0);
    const OctreeMesh& m = *mesh_;
    for (size_t s = 0; s < station_cell_.size(); ++s) {
        const int cell = station_cell_[s];
        if (cell < 0) continue;
        const auto& w = wr[s];
        const StationHCache& hc = cache[s];
        const cplx det = hc.det_H;
        if (std::abs(det) < 1e-30) continue;

        const cplx sx1 = (w[0] * hc.Hy2 - w[1] * hc.Hx2) / det;
        const cplx sy1 = (w[2] * hc.Hy2 - w[3] * hc.Hx2) / det;
        const cplx sx2 = (-w[0] * hc.Hy1 + w[1] * hc.Hx1) / det;
        const cplx sy2 = (-w[2] * hc.Hy1 + w[3] * hc.Hx1) / det;
        const cplx aHx1 = -(hc.Zxx * sx1 + hc.Zyx * sy1);
        const cplx aHy1 = -(hc.Zxy * sx1 + hc.Zyy * sy1);
        const cplx aHx2 = -(hc.Zxx * sx2 + hc.Zyx * sy2);
        const cplx aHy2 = -(hc.Zxy * sx2 + hc.Zyy * sy2);

        double cx, cy, cz;
        m.cell_center(cell, cx, cy, cz);
        CellGroups g;
        for (const CellEdge& ce : grid_.cell_edges[cell]) {
            const int e = ce.edge_id;
            const int d = grid_.edge_dir[e];
            if (d == 0) { if (grid_.edge_z[e] > cz) ++g.nxt; else ++g.nxb; }
            else if (d == 1) { if (grid_.edge_z[e] > cz) ++g.nyt; else ++g.nyb; }
            else {
                if (grid_.edge_x[e] > cx) ++g.nzpx; else ++g.nzmx;
                if (grid_.edge_y[e] > cy) ++g.nzpy; else ++g.nzmy;
            }
        }
        double wxt = g.nxt > 0 ? 0.5 / g.nxt : 0.0, wxb = g.nxb > 0 ? 0.5 / g.nxb : 0.0;
        double wyt = g.nyt > 0 ? 0.5 / g.nyt : 0.0, wyb = g.nyb > 0 ? 0.5 / g.nyb : 0.0;
        if (g.nxt == 0) wxb = g.nxb > 0 ? 1.0 / g.nxb : 0.0;
        if (g.nxb == 0) wxt = g.nxt > 0 ? 1.0 / g.nxt : 0.0;
        if (g.nyt == 0) wyb = g.nyb > 0 ? 1.0 / g.nyb : 0.0;
        if (g.nyb == 0) wyt = g.nyt > 0 ? 1.0 / g.nyt : 0.0;

        const cplx iz = cplx(1, 0) / (iwmu * hc.cdz);
        const cplx ix = cplx(1, 0) / (iwmu * hc.cdx);
        const cplx iy = cplx(1, 0) / (iwmu * hc.cdy);

        for (const CellEdge& ce : grid_.cell_edges[cell]) {
            const int e = ce.edge_id;
            const int d = grid_.edge_dir[e];
            if (d == 0) {
                const bool top = grid_.edge_z[e] > cz;
                const double we = top ? wxt : wxb;
                add(e, sx1 * we, sx2 * we);
                if (top && g.nxt > 0) {
                    const cplx wh = iz / double(g.nxt);
                    add(e, aHy1 * wh, aHy2 * wh);
                } else if (!top && g.nxb > 0) {
                    const cplx wh = -iz / double(g.nxb);
                    add(e, aHy1 * wh, aHy2 * wh);
                }
            } else if (d == 1) {
                const bool top = grid_.edge_z[e] > cz;
                const double we = top ? wyt : wyb;
                add(e, sy1 * we, sy2 * we);
                if (top && g.nyt > 0) {
                    const cplx wh = -iz / double(g.nyt);
                    add(e, aHx1 * wh, aHx2 * wh);
                } else if (!top && g.nyb > 0) {
                    const cplx wh = iz / double(g.nyb);
                    add(e, aHx1 * wh, aHx2 * wh);
                }
            } else {
                const bool py = grid_.edge_y[e] > cy;
                if (py && g.nzpy > 0) {
                    const cplx wh = iy / double(g.nzpy);
                    add(e, aHx1 * wh, aHx2 * wh);
                } else if (!py && g.nzmy > 0) {
                    const cplx wh = -iy / double(g.nzmy);
                    add(e, aHx1 * wh, aHx2 * wh);
                }
                const bool px = grid_.edge_x[e] > cx;
                if (px && g.nzpx > 0) {
                    const cplx wh = -ix / double(g.nzpx);
                    add(e, aHy1 * wh, aHy2 * wh);
                } else if (!px && g.nzmx > 0) {
                    const cplx wh = ix / double(g.nzmx);
                    add(e, aHy1 * wh, aHy2 * wh);
                }
            }
        }
    }
}

void FVProblem::adjoint_rhs(double omega, const std::vector<StationHCache>& cache,
                            const std::vector<std::array<cplx, 4>>& wr, std::vector<cplx>& rhs1,
                            std::vector<cplx>& rhs2) const {
    const int ne = num_edges();
    rhs1.assign(ne, cplx(0, 0));
    rhs2.assign(ne, cplx(0, 0));
    adjoint_rhs_core(omega, cache, wr, [&](int e, const cplx& v1, const cplx& v2) {
        rhs1[e] += v1;
        rhs2[e] += v2;
    });
}

std::vector<int> FVProblem::station_edges() const {
    std::vector<int> out;
    for (int cell : station_cell_)
        if (cell >= 0)
            for (const CellEdge& ce : grid_.cell_edges[cell]) out.push_back(ce.edge_id);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

void FVProblem::extract_impedance_compact(double omega, const std::vector<int>& edges, const cplx* e1,
                                          const cplx* e2, std::vector<std::array<cplx, 4>>& z,
                                          std::vector<StationHCache>& cache) const {
    auto pos = [&](int e) {
        const auto it = std::lower_bound(edges.begin(), edges.end(), e);
        if (it == edges.end() || *it != e) throw std::invalid_argument("edge outside the station edge list");
        return static_cast<size_t>(it - edges.begin());
    };
    auto g1 = [&](int e) -> const cplx& { return e1[pos(e)]; };
    auto g2 = [&](int e) -> const cplx& { return e2[pos(e)]; };
    extract_impedance_core(omega, g1, g2, z, cache);
}

void FVProblem::adjoint_rhs_compact(double omega, const std::vector<StationHCache>& cache,
                                    const std::vector<std::array<cplx, 4>>& wr, const std::vector<int>& edges,
                                    cplx* rhs1, cplx* rhs2) const {
    for (size_t k = 0; k < edges.size(); ++k) rhs1[k] = rhs2[k] = cplx(0, 0);
    adjoint_rhs_core(omega, cache, wr, [&](int e, const cplx& v1, const cplx& v2) {
        const auto it = std::lower_bound(edges.begin(), edges.end(), e);
        if (it == edges.end() || *it != e) throw std::invalid_argument("edge outside the station edge list");
        const size_t k = static_cast<size_t>(it - edges.begin());
        rhs1[k] += v1;
        rhs2[k] += v2;
    });
}

void FVProblem::station_table(StationTable& t) const {
    const OctreeMesh& m = *mesh_;
    const int ns = static_cast<int>(station_cell_.size());
    t = StationTable{};
    t.valid.assign(ns, 0);
    t.edge_ptr.assign(ns + 1, 0);
    t.cell_size.assign(3 * size_t(ns), 0.0);
    struct Contribution {
        int edge, station, kind, axis_sign, n;
        double we;
    };
    std::vector<Contribution> cs;
    for (int s = 0; s < ns; ++s) {
        const int cell = station_cell_[s];
        if (cell >= 0) {
            t.valid[s] = 1;
            double cx, cy, cz, dx, dy, dz;
            m.cell_center(cell, cx, cy, cz);
            m.cell_size_xyz(cell, dx, dy, dz);
            t.cell_size[3 * s] = dx;
            t.cell_size[3 * s + 1] = dy;
            t.cell_size[3 * s + 2] = dz;
            CellGroups g;
            for (const CellEdge& ce : grid_.cell_edges[cell]) {
                const int e = ce.edge_id;
                const int d = grid_.edge_dir[e];
                unsigned char role = static_cast<unsigned char>(d);
                if (d == 0 || d == 1) {
                    const bool top = grid_.edge_z[e] > cz;
                    if (top) role |= 4;
                    if (d == 0) { if (top) ++g.nxt; else ++g.nxb; }
                    else { if (top) ++g.nyt; else ++g.nyb; }
                } else {
                    const bool px = grid_.edge_x[e] > cx, py = grid_.edge_y[e] > cy;
                    if (px) role |= 4;
                    if (py) role |= 8;
                    if (px) ++g.nzpx; else ++g.nzmx;
                    if (py) ++g.nzpy; else ++g.nzmy;
                }
                t.edge.push_back(e);
                t.role.push_back(role);
            }
            // Weights and H terms exactly as adjoint_rhs_core forms them.
            double wxt = g.nxt > 0 ? 0.5 / g.nxt : 0.0, wxb = g.nxb > 0 ? 0.5 / g.nxb : 0.0;
            double wyt = g.nyt > 0 ? 0.5 / g.nyt : 0.0, wyb = g.nyb > 0 ? 0.5 / g.nyb : 0.0;
            if (g.nxt == 0) wxb = g.nxb > 0 ? 1.0 / g.nxb : 0.0;
            if (g.nxb == 0) wxt = g.nxt > 0 ? 1.0 / g.nxt : 0.0;
            if (g.nyt == 0) wyb = g.nyb > 0 ? 1.0 / g.nyb : 0.0;
            if (g.nyb == 0) wyt = g.nyt > 0 ? 1.0 / g.nyt : 0.0;
            for (const CellEdge& ce : grid_.cell_edges[cell]) {
                const int e = ce.edge_id;
                const int d = grid_.edge_dir[e];
                if (d == 0) {
                    const bool top = grid_.edge_z[e] > cz;
                    cs.push_back({e, s, 0, 0, 0, top ? wxt : wxb});
                    if (top && g.nxt > 0) cs.push_back({e, s, 2, 0, g.nxt, 0.0});
                    else if (!top && g.nxb > 0) cs.push_back({e, s, 2, 1, g.nxb, 0.0});
                } else if (d == 1) {
                    const bool top = grid_.edge_z[e] > cz;
                    cs.push_back({e, s, 1, 0, 0, top ? wyt : wyb});
                    if (top && g.nyt > 0) cs.push_back({e, s, 3, 1, g.nyt, 0.0});
                    else if (!top && g.nyb > 0) cs.push_back({e, s, 3, 0, g.nyb, 0.0});
                } else {
                    const bool py = grid_.edge_y[e] > cy;
                    if (py && g.nzpy > 0) cs.push_back({e, s, 3, 4, g.nzpy, 0.0});
                    else if (!py && g.nzmy > 0) cs.push_back({e, s, 3, 5, g.nzmy, 0.0});
                    const bool px = grid_.edge_x[e] > cx;
                    if (px && g.nzpx > 0) cs.push_back({e, s, 2, 3, g.nzpx, 0.0});
                    else if (!px && g.nzmx > 0) cs.push_back({e, s, 2, 2, g.nzmx, 0.0});
                }
            }
        }
        t.edge_ptr[s + 1] = static_cast<int>(t.edge.size());
    }
    t.uedge = station_edges();
    const int nu = static_cast<int>(t.uedge.size());
    t.contrib_ptr.assign(nu + 1, 0);
    auto pos = [&](int e) {
        return static_cast<int>(std::lower_bound(t.uedge.begin(), t.uedge.end(), e) - t.uedge.begin());
    };
    for (const Contribution& c : cs) ++t.contrib_ptr[pos(c.edge) + 1];
    for (int k = 0; k < nu; ++k) t.contrib_ptr[k + 1] += t.contrib_ptr[k];
    t.contrib.assign(4 * cs.size(), 0);
    t.contrib_we.assign(cs.size(), 0.0);
    std::vector<int> fill(t.contrib_ptr.begin(), t.contrib_ptr.end() - 1);
    for (const Contribution& c : cs) {  // stable: keeps the order adjoint_rhs adds in
        const int j = fill[pos(c.edge)]++;
        t.contrib[4 * j] = c.station;
        t.contrib[4 * j + 1] = c.kind;
        t.contrib[4 * j + 2] = c.axis_sign;
        t.contrib[4 * j + 3] = c.n;
        t.contrib_we[j] = c.we;
    }
}

void FVProblem::station_sigma(std::vector<double>& out) const {
    out.assign(station_cell_.size(), 0.0);
    for (size_t s = 0; s < station_cell_.size(); ++s)
        if (station_cell_[s] >= 0) out[s] = sigma[station_cell_[s]];
}

// g_c = Re(i omega sum_{e in c, unconstrained} lambda_e E_e sigma_c A_dual L / n_adj)
// with unconjugated lambda (:989-1059).
void FVProblem::sensitivity_weights(std::vector<double>& w) const {
    const int ne = num_edges(), nc = num_cells();
    w.assign(ne, 0.0);
    for (int e = 0; e < ne; ++e) {
        if (sens_constrained[e]) continue;
        int n_adj = 0;
        for (int a : grid_.edge_adj[e])
            if (a >= 0 && a < nc) ++n_adj;
        if (n_adj == 0) continue;
        w[e] = ops_.edge_dual_area(e) * grid_.edge_len[e] / n_adj;
    }
}

void FVProblem::sensitivity(double omega, const std::vector<cplx>& e_bg,
                            const std::vector<cplx>& lambda, std::vector<double>& sens) const {
    const int nc = num_cells();
    sens.assign(nc, 0.0);
    const cplx iw(0.0, omega);
    for (int c = 0; c < nc; ++c) {
        cplx s(0, 0);
        const double sc = sigma[c];
        for (const CellEdge& ce : grid_.cell_edges[c]) {
            const int e = ce.edge_id;
            if (sens_constrained[e]) continue;
            int n_adj = 0;
            for (int a : grid_.edge_adj[e])
                if (a >= 0 && a < nc) ++n_adj;
            if (n_adj == 0) continue;
            const double dme = sc * ops_.edge_dual_area(e) * grid_.edge_len[e] / n_adj;
            s += lambda[e] * e_bg[e] * dme;
        }
        sens[c] = (iw * s).real();
    }
}

// delta b = -i omega sum_c dMe/dsigma_c delta_sigma_c E_bg with the harmonic
// mean and L^3 volume of :945-984 (kept as the reference has it).
void FVProblem::perturbation_rhs(double omega, const std::vector<cplx>& e_bg,
                                 const std::vector<double>& dsig, std::vector<cplx>& rhs) const {
    const int ne = num_edges(), nc = num_cells();
    rhs.assign(ne, cplx(0, 0));
    const cplx iw(0.0, omega);
    for (int c = 0; c < nc; ++c) {
        if (std::abs(dsig[c]) < 1e-30) continue;
        const double sc = sigma[c];
        for (const CellEdge& ce : grid_.cell_edges[c]) {
            const int e = ce.edge_id;
            const double L = grid_.edge_len[e];
            const double vol = L * L * L;
            double sum_inv = 0.0;
            int n_adj = 0;
            for (int a : grid_.edge_adj[e]) {
                if (a >= 0 && a < nc && sigma[a] > 1e-30) {
                    sum_inv += 1.0 / sigma[a];
                    ++n_adj;
                }
            }
            const double savg = (n_adj > 0 && sum_inv > 1e-30) ? n_adj / sum_inv : 1e-8;
            const double dme = vol * savg * savg / (n_adj * sc * sc);
            rhs[e] -= iw * cplx(dsig[c] * dme, 0) * e_bg[e];
        }
    }
}

template <typename Get>
void FVProblem::delta_impedance_core(Get de, int polarization, const std::vector<StationHCache>& cache,
                                     std::vector<std::array<cplx, 4>>& dz) const {
    const int ns = static_cast<int>(station_cell_.size());
    dz.assign(ns, {cplx(0, 0), cplx(0, 0), cplx(0, 0), cplx(0, 0)});
    for (int s = 0; s < ns; ++s) {
        const int cell = station_cell_[s];
        if (cell < 0) continue;
        cplx dex, dey;
        int nx = 0, ny = 0;
        for (const CellEdge& ce : grid_.cell_edges[cell]) {
            const int e = ce.edge_id;
            if (grid_.edge_dir[e] == 0) { dex += de(e); ++nx; }
            else if (grid_.edge_dir[e] == 1) { dey += de(e); ++ny; }
        }
        if (nx > 0) dex /= cplx(nx, 0);
        if (ny > 0) dey /= cplx(ny, 0);
        const StationHCache& hc = cache[s];
        if (std::abs(hc.det_H) < 1e-30) continue;
        const cplx inv = cplx(1, 0) / hc.det_H;
        if (polarization == 1) {
            dz[s] = {dex * hc.Hy2 * inv, dex * (-hc.Hx2) * inv, dey * hc.Hy2 * inv, dey * (-hc.Hx2) * inv};
        } else {
            dz[s] = {dex * (-hc.Hy1) * inv, dex * hc.Hx1 * inv, dey * (-hc.Hy1) * inv, dey * hc.Hx1 * inv};
        }
    }
}

void FVProblem::delta_impedance(const std::vector<cplx>& de, int polarization,
                                const std::vector<StationHCache>& cache,
                                std::vector<std::array<cplx, 4>>& dz) const {
    delta_impedance_core([&](int e) -> const cplx& { return de[e]; }, polarization, cache, dz);
}

void FVProblem::delta_impedance_compact(const std::vector<int>& edges, const cplx* de, int polarization,
                                        const std::vector<StationHCache>& cache,
                                        std::vector<std::array<cplx, 4>>& dz) const {
    delta_impedance_core(
        [&](int e) -> const cplx& {
            const auto it = std::lower_bound(edges.begin(), edges.end(), e);
            if (it == edges.end() || *it != e) throw std::invalid_argument("edge outside the station edge list");
            return de[it - edges.begin()];
        },
        polarization, cache, dz);
}

}  // namespace m3d
