// extern "C" surface of the host side (include/m3d_host.h).
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "edi_io.h"
#include "forward_solver_gpu.h"
#include "m3d_host.h"

using namespace m3d;

namespace {

thread_local std::string g_host_error;

struct Array {
    int dtype = 0;
    std::vector<double> f64;
    std::vector<int> i32;
    std::vector<unsigned char> u8;
    long long count() const {
        return dtype == 0 ? (long long)f64.size() : dtype == 1 ? (long long)i32.size()
                          : dtype == 2 ? (long long)u8.size() : (long long)f64.size() / 2;
    }
    const void* data() const {
        return dtype == 1 ? (const void*)i32.data() : dtype == 2 ? (const void*)u8.data() : (const void*)f64.data();
    }
};

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_host_error = e.what();
        return 1;
    }
}

Array f64(const std::vector<double>& v) { Array a; a.dtype = 0; a.f64 = v; return a; }
Array i32(const std::vector<int>& v) { Array a; a.dtype = 1; a.i32 = v; return a; }
template <typename T>
Array u8(const std::vector<T>& v) {
    Array a; a.dtype = 2; a.u8.assign(v.begin(), v.end()); return a;
}
template <typename T>
void flatten(const std::vector<std::vector<T>>& vv, std::vector<int>& ptr, std::vector<T>& out) {
    ptr.assign(vv.size() + 1, 0);
    out.clear();
    for (size_t i = 0; i < vv.size(); ++i) {
        out.insert(out.end(), vv[i].begin(), vv[i].end());
        ptr[i + 1] = static_cast<int>(out.size());
    }
}

}  // namespace

struct m3dh_model {
    OctreeMesh mesh;
    ForwardSolverGPU fwd;
    std::map<std::string, Array> cache;
    std::vector<std::array<cplx, 4>> last_z;
    std::vector<StationHCache> hcache;
    std::vector<int> station_edges;  // sorted station-cell edges (m3dh_station_edges)

    Array build(const std::string& name) {
        const FVProblem& P = fwd.problem();
        const StaggeredGrid& g = P.grid();
        const DiscreteOperators& o = P.ops();
        const int nc = mesh.num_cells();
        if (name == "cell_cx" || name == "cell_cy" || name == "cell_cz" || name == "cell_size") {
            std::vector<double> v(nc);
            for (int c = 0; c < nc; ++c) {
                double x, y, z;
                mesh.cell_center(c, x, y, z);
                v[c] = name == "cell_cx" ? x : name == "cell_cy" ? y : name == "cell_cz" ? z : mesh.cell_size(c);
            }
            return f64(v);
        }
        if (name == "cell_level") {
            std::vector<int> v(nc);
            for (int c = 0; c < nc; ++c) v[c] = mesh.cell_level(c);
            return i32(v);
        }
        if (name == "cell_type") {
            std::vector<unsigned char> v(nc);
            for (int c = 0; c < nc; ++c) v[c] = static_cast<unsigned char>(mesh.cell_type(c));
            return u8(v);
        }
        if (name == "ops_cell_cx") return f64(o.cell_cx());
        if (name == "ops_cell_cy") return f64(o.cell_cy());
        if (name == "ops_cell_cz") return f64(o.cell_cz());
        if (name == "edge_x") return f64(g.edge_x);
        if (name == "edge_y") return f64(g.edge_y);
        if (name == "edge_z") return f64(g.edge_z);
        if (name == "edge_len") return f64(g.edge_len);
        if (name == "edge_dir") return u8(g.edge_dir);
        if (name == "edge_adj_ptr" || name == "edge_adj") {
            std::vector<int> ptr, out;
            flatten(g.edge_adj, ptr, out);
            return i32(name == "edge_adj_ptr" ? ptr : out);
        }
        if (name == "face_x") return f64(g.face_x);
        if (name == "face_y") return f64(g.face_y);
        if (name == "face_z") return f64(g.face_z);
        if (name == "face_area") return f64(g.face_area);
        if (name == "face_normal") return u8(g.face_normal);
        if (name == "face_minus") return i32(g.face_minus);
        if (name == "face_plus") return i32(g.face_plus);
        if (name == "face_edges_ptr" || name == "face_edges") {
            std::vector<int> ptr, out;
            flatten(g.face_edges, ptr, out);
            return i32(name == "face_edges_ptr" ? ptr : out);
        }
        if (name == "hanging_faces") {
            std::vector<int> v;
            for (const auto& h : g.hanging_faces) {
                v.push_back(h.coarse_face);
                for (int f : h.fine_faces) v.push_back(f);
            }
            return i32(v);
        }
        if (name == "hanging_edges") {
            std::vector<int> v;
            for (const auto& h : g.hanging_edges) {
                v.push_back(h.coarse_edge);
                v.push_back(h.fine_edges[0]);
                v.push_back(h.fine_edges[1]);
            }
            return i32(v);
        }
        if (name == "sens_weight") {
            std::vector<double> w;
            P.sensitivity_weights(w);
            return f64(w);
        }
        if (name == "cell_edges_ptr" || name == "cell_edges") {
            std::vector<int> ptr(nc + 1, 0), out;
            for (int c = 0; c < nc; ++c) {
                for (const auto& ce : g.cell_edges[c]) out.push_back(ce.edge_id);
                ptr[c + 1] = static_cast<int>(out.size());
            }
            return i32(name == "cell_edges_ptr" ? ptr : out);
        }
        auto csr = [&](const CsrReal& m, const std::string& base) -> Array {
            if (name == base + "_ptr") return i32(m.ptr);
            if (name == base + "_col") return i32(m.col);
            return f64(m.val);
        };
        if (name.rfind("Ce_", 0) == 0) return csr(o.curl(), "Ce");
        if (name.rfind("D_", 0) == 0) return csr(o.divergence(), "D");
        if (name.rfind("G_", 0) == 0) return csr(o.gradient_node(), "G");
        if (name.rfind("S_", 0) == 0) return csr(o.stiffness(), "S");
        if (name.rfind("L_", 0) == 0 && name != "L_diag") return csr(P.L_bc, "L");
        if (name == "L_diag") return f64(P.L_diag);
        if (name.rfind("offdiag_", 0) == 0) return csr(P.offdiag, "offdiag");
        if (name == "diag_s") return f64(P.diag_s);
        if (name == "edge_dual_area") {
            std::vector<double> v(o.num_edges());
            for (int e = 0; e < o.num_edges(); ++e) v[e] = o.edge_dual_area(e);
            return f64(v);
        }
        if (name == "face_dual_len") {
            std::vector<double> v(o.num_faces());
            for (int f = 0; f < o.num_faces(); ++f) v[f] = o.face_dual_length(f);
            return f64(v);
        }
        if (name == "mf_inv") {
            std::vector<double> v;
            o.face_mass_inv(v);
            return f64(v);
        }
        if (name == "node_x") return f64(o.node_x());
        if (name == "node_y") return f64(o.node_y());
        if (name == "node_z") return f64(o.node_z());
        if (name == "dead_edge") {
            std::vector<unsigned char> v(o.num_edges());
            for (int e = 0; e < o.num_edges(); ++e) v[e] = o.is_dead_edge(e);
            return u8(v);
        }
        if (name == "dead_face") {
            std::vector<unsigned char> v(o.num_faces());
            for (int f = 0; f < o.num_faces(); ++f) v[f] = o.is_dead_face(f);
            return u8(v);
        }
        if (name == "is_bc") return u8(P.is_bc);
        if (name == "is_dead") return u8(P.is_dead);
        if (name == "is_constrained") return u8(P.is_constrained);
        if (name == "dc_bc_edge") return u8(P.dc_bc_edge);
        if (name == "dc_bc_node") return u8(P.dc_bc_node);
        if (name == "sens_constrained") return u8(P.sens_constrained);
        if (name == "sigma") return f64(P.sigma);
        if (name == "me") return f64(P.me);
        if (name == "me_s") return f64(P.me_s);
        if (name == "station_cell") return i32(P.station_cells());
        if (name.rfind("st_", 0) == 0) {
            if (name == "st_sigma") {
                std::vector<double> v;
                P.station_sigma(v);
                return f64(v);
            }
            StationTable t;
            P.station_table(t);
            if (name == "st_valid") return i32(t.valid);
            if (name == "st_edge_ptr") return i32(t.edge_ptr);
            if (name == "st_edge") return i32(t.edge);
            if (name == "st_role") return u8(t.role);
            if (name == "st_cell_size") return f64(t.cell_size);
            if (name == "st_uedge") return i32(t.uedge);
            if (name == "st_contrib_ptr") return i32(t.contrib_ptr);
            if (name == "st_contrib") return i32(t.contrib);
            if (name == "st_contrib_we") return f64(t.contrib_we);
        }
        if (name == "hanging_pairs") {
            std::vector<int> v;
            for (const auto& h : o.hanging_pairs()) {
                v.push_back(h.coarse); v.push_back(h.child1); v.push_back(h.child2);
            }
            return i32(v);
        }
        if (name == "device_perm" || name == "device_colour_start") {
            m3d_operator* op = fwd.device_operator();
            if (!op) throw std::runtime_error("model has no device operator");
            int npad = 0, nc2 = 0;
            m3d_operator_info(op, &npad, &nc2, nullptr, nullptr);
            std::vector<int> perm(npad), cs(nc2 + 1);
            m3d_operator_permutation(op, perm.data(), cs.data());
            return i32(name == "device_perm" ? perm : cs);
        }
        throw std::invalid_argument("unknown array name: " + name);
    }
};

extern "C" {

const char* m3dh_last_error(void) { return g_host_error.c_str(); }

int m3dh_model_create(const m3dh_mesh_params* mp, const double* st_xyz, int n_st, const double* freqs,
                      int n_freqs, const m3dh_forward_params* fp, m3d_context* ctx, m3dh_model** out) {
    return guarded([&] {
        if (!mp || !out) throw std::invalid_argument("null argument");
        std::unique_ptr<m3dh_model> m(new m3dh_model);
        RefinementParams rp;
        rp.domain_x_min = mp->domain[0]; rp.domain_x_max = mp->domain[1];
        rp.domain_y_min = mp->domain[2]; rp.domain_y_max = mp->domain[3];
        rp.domain_z_min = mp->domain[4]; rp.domain_z_max = mp->domain[5];
        rp.min_level = mp->min_level; rp.max_level = mp->max_level;
        rp.station_refine_radius = mp->station_refine_radius;
        rp.station_refine_level = mp->station_refine_level;
        rp.sigma_bg = mp->sigma_bg;
        rp.max_interest_depth = mp->max_interest_depth;
        rp.air_surface_thickness = mp->air_surface_thickness;
        rp.air_mid_thickness = mp->air_mid_thickness;
        rp.air_upper_thickness = mp->air_upper_thickness;
        rp.transition_padding = mp->transition_padding;
        rp.station_x_min = mp->station_bbox[0]; rp.station_x_max = mp->station_bbox[1];
        rp.station_y_min = mp->station_bbox[2]; rp.station_y_max = mp->station_bbox[3];
        for (int r = 0; r < mp->n_regions; ++r) {
            const double* q = mp->regions + 8 * r;
            rp.refine_regions.push_back({q[0], q[1], q[2], q[3], q[4], q[5], int(q[6]), q[7]});
        }
        std::vector<std::array<double, 3>> st;
        for (int i = 0; i < n_st; ++i) st.push_back({st_xyz[3 * i], st_xyz[3 * i + 1], st_xyz[3 * i + 2]});
        std::vector<double> f(freqs, freqs + n_freqs);
        m->mesh.setup(rp, st, f);
        if (mp->flat_terrain) m->mesh.set_terrain([](double, double) { return 0.0; });
        ForwardParams p;
        if (fp) {
            p.bicgstab_tol = fp->bicgstab_tol;
            p.bicgstab_maxiter = fp->bicgstab_maxiter;
            p.air_z_threshold = fp->air_z_threshold;
            p.div_correction = fp->div_correction != 0;
            p.div_corr_iters = fp->div_corr_iters;
            p.div_corr_tol = fp->div_corr_tol;
            p.divcorr_outer_max = fp->divcorr_outer_max;
            p.divcorr_iter_per_dc = fp->divcorr_iter_per_dc;
        }
        m->fwd.setup(m->mesh, p, ctx);
        *out = m.release();
    });
}

int m3dh_model_destroy(m3dh_model* m) {
    return guarded([&] { delete m; });
}

int m3dh_get_array(m3dh_model* m, const char* name, int* dtype, long long* count, const void** data) {
    return guarded([&] {
        if (!m || !name) throw std::invalid_argument("null argument");
        auto it = m->cache.find(name);
        if (it == m->cache.end()) it = m->cache.emplace(name, m->build(name)).first;
        if (dtype) *dtype = it->second.dtype;
        if (count) *count = it->second.count();
        if (data) *data = it->second.data();
    });
}

long long m3dh_get_count(m3dh_model* m, const char* name) {
    if (!m || !name) return -1;
    const std::string n(name);
    const FVProblem& P = m->fwd.problem();
    if (n == "cells") return m->mesh.num_cells();
    if (n == "edges") return P.num_edges();
    if (n == "faces") return P.ops().num_faces();
    if (n == "nodes") return P.num_nodes();
    if (n == "hanging_faces") return (long long)P.grid().hanging_faces.size();
    if (n == "hanging_edges") return (long long)P.grid().hanging_edges.size();
    if (n == "stations") return (long long)P.station_cells().size();
    if (n == "colours") {
        m3d_operator* op = m->fwd.device_operator();
        int nc = 0;
        if (op) m3d_operator_info(op, nullptr, &nc, nullptr, nullptr);
        return nc;
    }
    return -1;
}

int m3dh_set_sigma(m3dh_model* m, const double* sigma) {
    return guarded([&] {
        const int nc = m->mesh.num_cells();
        m->fwd.set_sigma(std::vector<double>(sigma, sigma + nc));
        m->cache.clear();
    });
}

int m3dh_update_sigma_device(m3dh_model* m, const double* sigma) {
    return guarded([&] {
        const int nc = m->mesh.num_cells();
        m->fwd.update_sigma_device(std::vector<double>(sigma, sigma + nc));
        m->cache.clear();
    });
}

int m3dh_find_stations(m3dh_model* m, const double* xy, int n) {
    return guarded([&] {
        std::vector<std::array<double, 2>> v;
        for (int i = 0; i < n; ++i) v.push_back({xy[2 * i], xy[2 * i + 1]});
        m->fwd.find_stations(v);
        m->cache.clear();
        m->station_edges.clear();
    });
}

int m3dh_primary_field(m3dh_model* m, double omega, int dir, double* out) {
    return guarded([&] {
        std::vector<cplx> v;
        m->fwd.problem().primary_field(omega, dir, v);
        std::memcpy(out, v.data(), v.size() * sizeof(cplx));
    });
}

int m3dh_bc_values(m3dh_model* m, double omega, int dir, double* out) {
    return guarded([&] {
        std::vector<cplx> v;
        m->fwd.problem().bc_values(omega, dir, v);
        std::memcpy(out, v.data(), v.size() * sizeof(cplx));
    });
}

int m3dh_rhs_total(m3dh_model* m, const double* ebc, double* out) {
    return guarded([&] {
        const int n = m->fwd.problem().num_edges();
        std::vector<cplx> e(reinterpret_cast<const cplx*>(ebc), reinterpret_cast<const cplx*>(ebc) + n), v;
        m->fwd.problem().rhs_total(e, v);
        std::memcpy(out, v.data(), v.size() * sizeof(cplx));
    });
}

int m3dh_phi0(m3dh_model* m, double omega, const double* rhs, double* out) {
    return guarded([&] {
        const int n = m->fwd.problem().num_edges();
        std::vector<cplx> r(reinterpret_cast<const cplx*>(rhs), reinterpret_cast<const cplx*>(rhs) + n), v;
        m->fwd.problem().phi0(omega, r, v);
        std::memcpy(out, v.data(), v.size() * sizeof(cplx));
    });
}

int m3dh_extract_impedance(m3dh_model* m, double omega, const double* e1, const double* e2, double* z_out) {
    return guarded([&] {
        const int n = m->fwd.problem().num_edges();
        const cplx* a = reinterpret_cast<const cplx*>(e1);
        const cplx* b = reinterpret_cast<const cplx*>(e2);
        std::vector<cplx> v1(a, a + n), v2(b, b + n);
        m->fwd.problem().extract_impedance(omega, v1, v2, m->last_z, m->hcache);
        m->fwd.set_station_cache(m->hcache);
        std::memcpy(z_out, static_cast<const void*>(m->last_z.data()), m->last_z.size() * sizeof(std::array<cplx, 4>));
    });
}

int m3dh_adjoint_rhs(m3dh_model* m, double omega, const double* wr, double* rhs1, double* rhs2) {
    return guarded([&] {
        const size_t ns = m->fwd.problem().station_cells().size();
        std::vector<std::array<cplx, 4>> w(ns);
        std::memcpy(static_cast<void*>(w.data()), wr, ns * sizeof(std::array<cplx, 4>));
        std::vector<cplx> r1, r2;
        m->fwd.problem().adjoint_rhs(omega, m->hcache, w, r1, r2);
        std::memcpy(rhs1, r1.data(), r1.size() * sizeof(cplx));
        std::memcpy(rhs2, r2.data(), r2.size() * sizeof(cplx));
    });
}

int m3dh_station_edges(m3dh_model* m, int* count, const int** edges) {
    return guarded([&] {
        if (!count || !edges) throw std::invalid_argument("null argument");
        m->station_edges = m->fwd.problem().station_edges();
        *count = static_cast<int>(m->station_edges.size());
        *edges = m->station_edges.data();
    });
}

int m3dh_extract_impedance_compact(m3dh_model* m, double omega, const double* e1c, const double* e2c,
                                   double* z_out) {
    return guarded([&] {
        if (m->station_edges.empty()) m->station_edges = m->fwd.problem().station_edges();
        m->fwd.problem().extract_impedance_compact(omega, m->station_edges, reinterpret_cast<const cplx*>(e1c),
                                                   reinterpret_cast<const cplx*>(e2c), m->last_z, m->hcache);
        m->fwd.set_station_cache(m->hcache);
        std::memcpy(z_out, static_cast<const void*>(m->last_z.data()), m->last_z.size() * sizeof(std::array<cplx, 4>));
    });
}

int m3dh_adjoint_rhs_compact(m3dh_model* m, double omega, const double* wr, double* rhs1c, double* rhs2c) {
    return guarded([&] {
        if (m->station_edges.empty()) m->station_edges = m->fwd.problem().station_edges();
        const size_t ns = m->fwd.problem().station_cells().size();
        std::vector<std::array<cplx, 4>> w(ns);
        std::memcpy(static_cast<void*>(w.data()), wr, ns * sizeof(std::array<cplx, 4>));
        m->fwd.problem().adjoint_rhs_compact(omega, m->hcache, w, m->station_edges, reinterpret_cast<cplx*>(rhs1c),
                                             reinterpret_cast<cplx*>(rhs2c));
    });
}

int m3dh_perturbation_rhs(m3dh_model* m, double omega, const double* e_bg, const double* dsigma,
                          double* rhs) {
    return guarded([&] {
        const int n = m->fwd.problem().num_edges(), nc = m->fwd.problem().num_cells();
        const cplx* a = reinterpret_cast<const cplx*>(e_bg);
        std::vector<cplx> e(a, a + n), r;
        m->fwd.problem().perturbation_rhs(omega, e, std::vector<double>(dsigma, dsigma + nc), r);
        std::memcpy(rhs, r.data(), r.size() * sizeof(cplx));
    });
}

int m3dh_delta_impedance_compact(m3dh_model* m, const double* dec, int polarization, double* dz_out) {
    return guarded([&] {
        if (polarization != 1 && polarization != 2) throw std::invalid_argument("polarization must be 1 or 2");
        if (m->station_edges.empty()) m->station_edges = m->fwd.problem().station_edges();
        std::vector<std::array<cplx, 4>> dz;
        m->fwd.problem().delta_impedance_compact(m->station_edges, reinterpret_cast<const cplx*>(dec),
                                                 polarization, m->hcache, dz);
        std::memcpy(dz_out, static_cast<const void*>(dz.data()), dz.size() * sizeof(std::array<cplx, 4>));
    });
}

int m3dh_sensitivity(m3dh_model* m, double omega, const double* e_bg, const double* lambda, double* sens) {
    return guarded([&] {
        const int n = m->fwd.problem().num_edges();
        const cplx* a = reinterpret_cast<const cplx*>(e_bg);
        const cplx* b = reinterpret_cast<const cplx*>(lambda);
        std::vector<cplx> e(a, a + n), l(b, b + n);
        std::vector<double> s;
        m->fwd.problem().sensitivity(omega, e, l, s);
        std::memcpy(sens, s.data(), s.size() * sizeof(double));
    });
}

m3d_operator* m3dh_device_operator(m3dh_model* m) { return m ? m->fwd.device_operator() : nullptr; }

int m3dh_solve_frequencies(m3dh_model* m, const double* freqs, int nf, const m3d_solve_params* prm,
                           double* e_out, double* z_out, m3d_solve_result* results) {
    return guarded([&] {
        m3d_operator* op = m->fwd.device_operator();
        if (!op) throw std::runtime_error("model has no device operator");
        const FVProblem& P = m->fwd.problem();
        const int n = P.num_edges();
        const int ns = 2 * nf;
        std::vector<double> rhs(2 * size_t(ns) * n), aux(2 * size_t(ns) * n);
        std::vector<double> om(ns);
        std::vector<cplx> ebc, rt, ep;
        for (int f = 0; f < nf; ++f) {
            const double w = kTwoPi * freqs[f];
            for (int pol = 0; pol < 2; ++pol) {
                const int s = 2 * f + pol;
                om[s] = w;
                P.bc_values(w, pol, ebc);
                P.rhs_total(ebc, rt);
                P.primary_field(w, pol, ep);
                std::memcpy(&rhs[2 * size_t(s) * n], rt.data(), sizeof(cplx) * n);
                std::memcpy(&aux[2 * size_t(s) * n], ep.data(), sizeof(cplx) * n);
            }
        }
        std::memset(e_out, 0, sizeof(double) * 2 * size_t(ns) * n);
        if (m3d_solve_batch(op, ns, M3D_MODE_FORWARD, om.data(), rhs.data(), aux.data(), e_out, prm, results) != M3D_OK)
            throw std::runtime_error(m3d_last_error());
        if (z_out) {
            const size_t nst = P.station_cells().size();
            for (int f = 0; f < nf; ++f) {
                const cplx* a = reinterpret_cast<const cplx*>(e_out) + size_t(2 * f) * n;
                const cplx* b = reinterpret_cast<const cplx*>(e_out) + size_t(2 * f + 1) * n;
                std::vector<cplx> v1(a, a + n), v2(b, b + n);
                std::vector<std::array<cplx, 4>> z;
                P.extract_impedance(om[2 * f], v1, v2, z, m->hcache);
                std::memcpy(z_out + 8 * nst * size_t(f), static_cast<const void*>(z.data()), nst * sizeof(std::array<cplx, 4>));
            }
            m->fwd.set_station_cache(m->hcache);
        }
    });
}

int m3dh_adjoint_solve(m3dh_model* m, const double* omegas, int nsys, const double* rhs,
                       const m3d_solve_params* prm, double* x_out, m3d_solve_result* results) {
    return guarded([&] {
        m3d_operator* op = m->fwd.device_operator();
        if (!op) throw std::runtime_error("model has no device operator");
        const int n = m->fwd.problem().num_edges();
        std::memset(x_out, 0, sizeof(double) * 2 * size_t(nsys) * n);
        if (m3d_solve_batch(op, nsys, M3D_MODE_INTERLEAVED, omegas, rhs, nullptr, x_out, prm, results) != M3D_OK)
            throw std::runtime_error(m3d_last_error());
    });
}


// EDI files ----------------------------------------------------------------

namespace {

using m3d::data::Datum;
using m3d::data::MTResponse;

Datum* component(MTResponse& r, int k) {
    Datum* d[6] = {&r.Zxx, &r.Zxy, &r.Zyx, &r.Zyy, &r.Tx, &r.Ty};
    return d[k];
}

// Responses of one station at one frequency to / from the flat arrays:
// z [4] complex, z_err [4], z_w [4], t [2] complex, t_err [2], t_w [2].
void put_response(const MTResponse& r0, size_t i, double* z, double* ze, double* zw, double* t, double* te,
                  double* tw) {
    MTResponse r = r0;
    for (int k = 0; k < 4; ++k) {
        const Datum* d = component(r, k);
        if (z) { z[2 * (4 * i + k)] = d->value.real(); z[2 * (4 * i + k) + 1] = d->value.imag(); }
        if (ze) ze[4 * i + k] = d->error;
        if (zw) zw[4 * i + k] = d->weight;
    }
    for (int k = 0; k < 2; ++k) {
        const Datum* d = component(r, 4 + k);
        if (t) { t[2 * (2 * i + k)] = d->value.real(); t[2 * (2 * i + k) + 1] = d->value.imag(); }
        if (te) te[2 * i + k] = d->error;
        if (tw) tw[2 * i + k] = d->weight;
    }
}

MTResponse get_response(size_t i, const double* z, const double* ze, const double* zw, const double* t,
                        const double* te, const double* tw) {
    MTResponse r;
    for (int k = 0; k < 4; ++k) {
        Datum* d = component(r, k);
        d->value = {z[2 * (4 * i + k)], z[2 * (4 * i + k) + 1]};
        d->error = ze[4 * i + k];
        d->weight = zw[4 * i + k];
    }
    for (int k = 0; k < 2; ++k) {
        Datum* d = component(r, 4 + k);
        if (t) d->value = {t[2 * (2 * i + k)], t[2 * (2 * i + k) + 1]};
        if (te) d->error = te[2 * i + k];
        d->weight = tw ? tw[2 * i + k] : 0.0;
    }
    return r;
}

void put_station(const m3d::data::Station& st, size_t s, char* names, int name_len, double* geo) {
    if (names && name_len > 0) {
        char* dst = names + s * size_t(name_len);
        std::memset(dst, 0, size_t(name_len));
        std::strncpy(dst, st.name.c_str(), size_t(name_len) - 1);
    }
    if (geo) {
        geo[4 * s] = st.lat; geo[4 * s + 1] = st.lon; geo[4 * s + 2] = st.z;
        geo[4 * s + 3] = st.has_geo ? 1.0 : 0.0;
    }
}

m3d::data::Station get_station(size_t s, const char* names, int name_len, const double* geo) {
    m3d::data::Station st;
    if (names && name_len > 0) {
        const char* src = names + s * size_t(name_len);
        st.name.assign(src, strnlen(src, size_t(name_len)));
    }
    if (geo) {
        st.lat = geo[4 * s]; st.lon = geo[4 * s + 1]; st.z = geo[4 * s + 2]; st.has_geo = geo[4 * s + 3] != 0.0;
    }
    return st;
}

}  // namespace

int m3dh_edi_read(const char* path, int* nfreq, char* name, int name_len, double* geo, double* freqs,
                  double* z, double* z_err, double* z_w, double* t, double* t_err, double* t_w) {
    return guarded([&] {
        if (!path || !nfreq) throw std::invalid_argument("null argument");
        const m3d::io::EDIRecord rec = m3d::io::read_edi(path);
        *nfreq = static_cast<int>(rec.frequencies.size());
        put_station(rec.station, 0, name, name_len, geo);
        for (size_t f = 0; f < rec.frequencies.size(); ++f) {
            if (freqs) freqs[f] = rec.frequencies[f];
            put_response(rec.responses[f], f, z, z_err, z_w, t, t_err, t_w);
        }
    });
}

int m3dh_edi_write(const char* path, const char* name, const double* geo, int nfreq, const double* freqs,
                   const double* z, const double* z_err, const double* z_w, const double* t, const double* t_err,
                   const double* t_w) {
    return guarded([&] {
        if (!path || nfreq < 0 || (nfreq > 0 && (!freqs || !z || !z_err || !z_w)))
            throw std::invalid_argument("null argument");
        m3d::data::Station st = get_station(0, name, name ? int(std::strlen(name)) + 1 : 0, geo);
        std::vector<MTResponse> r(nfreq);
        for (int f = 0; f < nfreq; ++f) r[f] = get_response(f, z, z_err, z_w, t, t_err, t_w);
        m3d::io::write_edi(path, st, std::vector<double>(freqs, freqs + nfreq), r);
    });
}

int m3dh_edi_load_directory(const char* dir, int* ns, int* nf, char* names, int name_len, double* geo,
                            double* freqs, double* z, double* z_err, double* z_w, double* t, double* t_err,
                            double* t_w) {
    return guarded([&] {
        if (!dir || !ns || !nf) throw std::invalid_argument("null argument");
        m3d::data::MTData d;
        m3d::io::load_edi_directory(dir, d);
        *ns = d.num_stations();
        *nf = d.num_frequencies();
        for (int s = 0; s < d.num_stations(); ++s) put_station(d.station(s), size_t(s), names, name_len, geo);
        for (int f = 0; f < d.num_frequencies(); ++f) {
            if (freqs) freqs[f] = d.frequencies()[f];
            for (int s = 0; s < d.num_stations(); ++s)
                put_response(d.observed(s, f), size_t(f) * d.num_stations() + s, z, z_err, z_w, t, t_err, t_w);
        }
    });
}

int m3dh_edi_save_directory(const char* dir, int ns, int nf, const char* names, int name_len, const double* geo,
                            const double* freqs, const double* z, const double* z_err, const double* z_w,
                            const double* t, const double* t_err, const double* t_w) {
    return guarded([&] {
        if (!dir || ns < 0 || nf < 0 || (nf > 0 && (!freqs || !z || !z_err || !z_w)))
            throw std::invalid_argument("null argument");
        m3d::data::MTData d;
        d.set_frequencies(std::vector<double>(freqs, freqs + nf));
        for (int s = 0; s < ns; ++s) d.add_station(get_station(size_t(s), names, name_len, geo));
        for (int f = 0; f < nf; ++f)
            for (int s = 0; s < ns; ++s)
                d.set_observed(s, f, get_response(size_t(f) * ns + s, z, z_err, z_w, t, t_err, t_w));
        m3d::io::save_edi_directory(dir, d);
    });
}

}  // extern "C"
