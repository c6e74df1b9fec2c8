// ForwardSolverFV-shaped driver over the device C-ABI; see forward_solver_gpu.h.
#include "forward_solver_gpu.h"

#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace m3d {

// The driver is a complete IForwardSolver: a caller holding the interface
// (InversionFV, ImplicitJtJOperatorFV) can be handed this class.
static_assert(std::is_base_of<forward::IForwardSolver, ForwardSolverGPU>::value, "IForwardSolver base");
static_assert(!std::is_abstract<ForwardSolverGPU>::value, "every IForwardSolver member is implemented");

ForwardSolverGPU::~ForwardSolverGPU() {
    if (op_) m3d_operator_destroy(op_);
}

void ForwardSolverGPU::check(int status) const {
    if (status != M3D_OK) throw std::runtime_error(std::string("device solve: ") + m3d_last_error());
}

m3d_solve_params ForwardSolverGPU::solve_params() const {
    const ForwardParams& p = prob_.params();
    m3d_solve_params s{};
    s.tol = p.bicgstab_tol;
    s.maxiter = p.bicgstab_maxiter;
    s.div_correction = p.div_correction ? 1 : 0;
    s.div_corr_iters = p.div_corr_iters;
    s.div_corr_tol = p.div_corr_tol;
    s.divcorr_outer_max = p.divcorr_outer_max;
    s.divcorr_iter_per_dc = p.divcorr_iter_per_dc;
    s.precision = M3D_PREC_FP64;
    s.inner_tol = 0.0;
    return s;
}

void ForwardSolverGPU::setup(const OctreeMesh& mesh, const ForwardParams& params, m3d_context* ctx) {
    prob_.setup(mesh, params);
    ctx_ = ctx;
    if (op_) {
        m3d_operator_destroy(op_);
        op_ = nullptr;
    }
    if (!ctx_) return;  // host-only use
    std::vector<int> tri;
    for (const auto& hp : prob_.ops().hanging_pairs()) {
        tri.push_back(hp.coarse);
        tri.push_back(hp.child1);
        tri.push_back(hp.child2);
    }
    std::vector<unsigned char> cons(prob_.is_constrained.begin(), prob_.is_constrained.end());
    std::vector<unsigned char> bc(prob_.is_bc.begin(), prob_.is_bc.end());
    m3d_operator_desc d{};
    d.n = prob_.num_edges();
    d.rowptr = prob_.offdiag.ptr.data();
    d.colidx = prob_.offdiag.col.data();
    d.values = prob_.offdiag.val.data();
    d.diag_s = prob_.diag_s.data();
    d.constrained = cons.data();
    d.bc = bc.data();
    d.n_hanging = static_cast<int>(tri.size() / 3);
    d.hanging = tri.empty() ? nullptr : tri.data();
    check(m3d_operator_create(ctx_, &d, &op_));
    // Edge-side data of the edge mass, for sigma updates on the device.
    const auto& ops = prob_.ops();
    const auto& adj = ops.edge_adjacency();
    std::vector<int> aptr(d.n + 1, 0), acell;
    for (int e = 0; e < d.n; ++e) {
        acell.insert(acell.end(), adj[e].begin(), adj[e].end());
        aptr[e + 1] = static_cast<int>(acell.size());
    }
    std::vector<unsigned char> massless(d.n, 0), dead(d.n, 0);
    for (const auto& hp : ops.hanging_pairs()) massless[hp.coarse] = 1;
    for (int e = 0; e < d.n; ++e) dead[e] = ops.is_dead_edge(e) ? 1 : 0;
    m3d_mass_structure ms{};
    ms.nc = prob_.num_cells();
    ms.adj_ptr = aptr.data();
    ms.adj_cell = acell.empty() ? nullptr : acell.data();
    ms.dual_area = ops.edge_dual_areas().data();
    ms.length = ops.edge_lengths().data();
    ms.massless = massless.data();
    ms.dead = dead.data();
    check(m3d_operator_set_mass_structure(op_, &ms));
}

void ForwardSolverGPU::update_sigma_device(const std::vector<double>& sigma) {
    if (static_cast<int>(sigma.size()) != prob_.num_cells())
        throw std::invalid_argument("sigma size does not match the number of cells");
    if (!op_) throw std::runtime_error("update_sigma_device needs a device operator");
    check(m3d_operator_update_sigma(op_, sigma.data()));
    prob_.sigma = sigma;
}

void ForwardSolverGPU::update_sigma(const model::ConductivityModel& model) {
    const int nc = prob_.num_cells();
    if (model.num_params() < nc) throw std::invalid_argument("conductivity model has fewer cells than the mesh");
    std::vector<double> sigma(nc);
    for (int c = 0; c < nc; ++c) sigma[c] = model.sigma(c);
    set_sigma(sigma);
}

void ForwardSolverGPU::set_sigma(const std::vector<double>& sigma) {
    prob_.set_sigma(sigma);
    if (!op_) return;
    check(m3d_operator_set_mass(op_, prob_.me.data()));
    const auto& G = prob_.ops().gradient_node();
    std::vector<unsigned char> bce(prob_.dc_bc_edge.begin(), prob_.dc_bc_edge.end());
    std::vector<unsigned char> dead(prob_.is_dead.begin(), prob_.is_dead.end());
    std::vector<unsigned char> bcn(prob_.dc_bc_node.begin(), prob_.dc_bc_node.end());
    m3d_divcorr_desc d{};
    d.nn = prob_.num_nodes();
    d.g_col = G.col.data();
    d.g_val = G.val.data();
    d.me_sigma = prob_.me_s.data();
    d.bc_edge = bce.data();
    d.dead_edge = dead.data();
    d.bc_node = bcn.data();
    d.L_rowptr = prob_.L_bc.ptr.data();
    d.L_colidx = prob_.L_bc.col.data();
    d.L_values = prob_.L_bc.val.data();
    d.L_diag = prob_.L_diag.data();
    check(m3d_operator_set_divcorr(op_, &d));
    // Sensitivity: cells -> edges, per-edge weights and sigma.
    const auto& grid = prob_.grid();
    std::vector<int> cptr(prob_.num_cells() + 1, 0), cedge;
    for (int c = 0; c < prob_.num_cells(); ++c) {
        for (const auto& ce : grid.cell_edges[c]) cedge.push_back(ce.edge_id);
        cptr[c + 1] = static_cast<int>(cedge.size());
    }
    std::vector<double> w;
    prob_.sensitivity_weights(w);
    m3d_sensitivity_desc sd{};
    sd.nc = prob_.num_cells();
    sd.cell_ptr = cptr.data();
    sd.cell_edges = cedge.data();
    sd.edge_weight = w.data();
    sd.sigma = prob_.sigma.data();
    check(m3d_operator_set_sensitivity(op_, &sd));
}

void ForwardSolverGPU::map_stations(const data::MTData& observed) {
    if (stations_mapped_) return;
    std::vector<std::array<double, 2>> xy;
    for (int s = 0; s < observed.num_stations(); ++s) xy.push_back({observed.station(s).x, observed.station(s).y});
    prob_.find_stations(xy);
    stations_mapped_ = true;
}

void ForwardSolverGPU::compute_responses(const data::MTData& observed, data::MTData& predicted) {
    map_stations(observed);
    const auto& f = observed.frequencies();
    if (freq_cb_)
        for (size_t fi = 0; fi < f.size(); ++fi) freq_cb_(int(fi), int(f.size()), 0.0, "Forward");
    std::vector<std::vector<cplx>> e;
    std::vector<std::vector<std::array<cplx, 4>>> z;
    solve_frequencies(f, e, z, &last_);
    for (size_t fi = 0; fi < f.size(); ++fi)
        for (size_t s = 0; s < z[fi].size(); ++s) {
            data::MTResponse r;
            r.Zxx.value = z[fi][s][0];
            r.Zxy.value = z[fi][s][1];
            r.Zyx.value = z[fi][s][2];
            r.Zyy.value = z[fi][s][3];
            predicted.set_predicted(int(s), int(fi), r);
        }
    // Keep the last frequency's fields and caches, as a sequential sweep would.
    if (!f.empty()) {
        e1_ = e[2 * (f.size() - 1)];
        e2_ = e[2 * (f.size() - 1) + 1];
        omega_ = kTwoPi * f.back();
        std::vector<std::array<cplx, 4>> zz;
        prob_.extract_impedance(omega_, e1_, e2_, zz, cache_);
    }
}

void ForwardSolverGPU::compute_single_frequency(int freq_idx, const data::MTData& observed,
                                                data::MTData& predicted) {
    map_stations(observed);
    solve_frequency(observed.frequencies()[freq_idx], freq_idx, observed, predicted);
}

void ForwardSolverGPU::solve_frequency(double freq_hz, int freq_idx, const data::MTData& /*observed*/,
                                       data::MTData& predicted) {
    std::vector<std::vector<cplx>> e;
    std::vector<std::vector<std::array<cplx, 4>>> z;
    solve_frequencies({freq_hz}, e, z, &last_);
    e1_ = e[0];
    e2_ = e[1];
    omega_ = kTwoPi * freq_hz;
    std::vector<std::array<cplx, 4>> zz;
    prob_.extract_impedance(omega_, e1_, e2_, zz, cache_);
    for (size_t s = 0; s < zz.size(); ++s) {
        data::MTResponse r;
        r.Zxx.value = zz[s][0];
        r.Zxy.value = zz[s][1];
        r.Zyx.value = zz[s][2];
        r.Zyy.value = zz[s][3];
        predicted.set_predicted(int(s), freq_idx, r);
    }
}

void ForwardSolverGPU::factorize_frequency(double freq_hz) { omega_ = kTwoPi * freq_hz; }

// The device operator is frequency independent and stays resident; nothing
// to release between frequencies.
void ForwardSolverGPU::release_factorization() {}

void ForwardSolverGPU::solve_frequencies(const std::vector<double>& freqs,
                                         std::vector<std::vector<cplx>>& e_out,
                                         std::vector<std::vector<std::array<cplx, 4>>>& z_out,
                                         std::vector<m3d_solve_result>* results) {
    if (!op_) throw std::runtime_error("ForwardSolverGPU has no device operator");
    const int n = prob_.num_edges();
    const int nf = static_cast<int>(freqs.size());
    const int ns = 2 * nf;
    std::vector<double> rhs(2 * size_t(ns) * n), aux(2 * size_t(ns) * n), x(2 * size_t(ns) * n, 0.0);
    std::vector<double> om(ns);
    std::vector<cplx> ebc, rt, ep;
    for (int f = 0; f < nf; ++f) {
        const double w = kTwoPi * freqs[f];
        for (int pol = 0; pol < 2; ++pol) {
            const int sys = 2 * f + pol;
            om[sys] = w;
            prob_.bc_values(w, pol, ebc);
            prob_.rhs_total(ebc, rt);
            prob_.primary_field(w, pol, ep);
            std::memcpy(&rhs[2 * size_t(sys) * n], rt.data(), sizeof(cplx) * n);
            std::memcpy(&aux[2 * size_t(sys) * n], ep.data(), sizeof(cplx) * n);
        }
    }
    m3d_solve_params sp = solve_params();
    std::vector<m3d_solve_result> res(ns);
    check(m3d_solve_batch(op_, ns, M3D_MODE_FORWARD, om.data(), rhs.data(), aux.data(), x.data(), &sp,
                          res.data()));
    e_out.assign(ns, std::vector<cplx>(n));
    for (int sys = 0; sys < ns; ++sys) std::memcpy(e_out[sys].data(), &x[2 * size_t(sys) * n], sizeof(cplx) * n);
    z_out.assign(nf, {});
    for (int f = 0; f < nf; ++f) {
        std::vector<StationHCache> c;
        prob_.extract_impedance(om[2 * f], e_out[2 * f], e_out[2 * f + 1], z_out[f], c);
    }
    if (results) *results = res;
}

void ForwardSolverGPU::build_adjoint_rhs_from_residual(int /*freq_idx*/,
                                                       const std::vector<std::array<cplx, 4>>& wr,
                                                       std::vector<cplx>& r1, std::vector<cplx>& r2) {
    prob_.adjoint_rhs(omega_, cache_, wr, r1, r2);
}

void ForwardSolverGPU::adjoint_solve_batch(const std::vector<double>& omegas,
                                           const std::vector<std::vector<cplx>>& rhs,
                                           std::vector<std::vector<cplx>>& lambda,
                                           std::vector<m3d_solve_result>* results) {
    if (!op_) throw std::runtime_error("ForwardSolverGPU has no device operator");
    const int n = prob_.num_edges();
    const int ns = static_cast<int>(rhs.size());
    std::vector<double> b(2 * size_t(ns) * n), x(2 * size_t(ns) * n, 0.0);
    for (int s = 0; s < ns; ++s) std::memcpy(&b[2 * size_t(s) * n], rhs[s].data(), sizeof(cplx) * n);
    m3d_solve_params sp = solve_params();
    if (tol_override_ > 0) sp.tol = tol_override_;
    std::vector<m3d_solve_result> res(ns);
    check(m3d_solve_batch(op_, ns, M3D_MODE_INTERLEAVED, omegas.data(), b.data(), nullptr, x.data(), &sp,
                          res.data()));
    lambda.assign(ns, std::vector<cplx>(n));
    for (int s = 0; s < ns; ++s) std::memcpy(lambda[s].data(), &x[2 * size_t(s) * n], sizeof(cplx) * n);
    if (results) *results = res;
}

void ForwardSolverGPU::adjoint_solve_complex(const std::vector<cplx>& rhs, std::vector<cplx>& lambda) {
    std::vector<std::vector<cplx>> out;
    adjoint_solve_batch({omega_}, {rhs}, out, &last_);
    lambda = std::move(out[0]);
}

void ForwardSolverGPU::solve_rhs_complex(const std::vector<cplx>& rhs, std::vector<cplx>& solution) {
    adjoint_solve_complex(rhs, solution);
}

void ForwardSolverGPU::compute_sensitivity_complex(const std::vector<cplx>& e_bg,
                                                   const std::vector<cplx>& lambda,
                                                   std::vector<double>& sens) {
    prob_.sensitivity(omega_, e_bg, lambda, sens);
}

// The reference picks E1_ for polarization == 1 and E2_ otherwise (:952).
void ForwardSolverGPU::build_perturbation_rhs_complex(int polarization,
                                                      const std::vector<double>& delta_sigma,
                                                      std::vector<cplx>& pert_rhs) {
    prob_.perturbation_rhs(omega_, polarization == 1 ? e1_ : e2_, delta_sigma, pert_rhs);
}

void ForwardSolverGPU::extract_delta_impedance_complex(const std::vector<cplx>& de, int polarization,
                                                       std::vector<std::array<cplx, 4>>& dz) {
    prob_.delta_impedance(de, polarization, cache_, dz);
}

}  // namespace m3d
