// Forward / adjoint solver with the public interface of ForwardSolverFV
// (include/maple3dmt/forward/forward_solver_fv.h:77-190, the IForwardSolver
// plugin of iforward_solver.h:31-115) whose solves run on the device through
// the C-ABI of include/m3d_device.h.
//
// Host work stays here (flags, right-hand sides, impedance, adjoint
// right-hand side, sensitivity); every Krylov / DivCorr solve is one batched
// device call.  Besides the reference methods it offers batched variants that
// solve several frequencies (and both polarisations) in one device call.
#pragma once

#include <array>
#include <complex>
#include <vector>

#include "fv_problem.h"
#include "iforward_solver.h"
#include "m3d_device.h"
#include "mt_data.h"
#include "octree_mesh.h"

namespace m3d {

class ForwardSolverGPU : public forward::IForwardSolver {
public:
    ForwardSolverGPU() = default;
    ~ForwardSolverGPU() override;
    ForwardSolverGPU(const ForwardSolverGPU&) = delete;
    ForwardSolverGPU& operator=(const ForwardSolverGPU&) = delete;

    // setup (forward_solver_fv.cpp:28-47): grid, operators, device operator.
    void setup(const OctreeMesh& mesh, const ForwardParams& params, m3d_context* ctx);
    // set_sigma / update_sigma (:49-58): host sigma data and the device mass
    // and DivCorr arrays.
    void set_sigma(const std::vector<double>& sigma);
    // update_sigma for a line-search trial: the device recomputes Me, the
    // DivCorr weights and Laplacian values and the sensitivity sigma itself
    // (m3d_operator_update_sigma; equal to set_sigma's device arrays).  The
    // host keeps the new sigma; its own Me and Laplacian copies are not
    // refreshed.  Needs one set_sigma first (the Laplacian pattern).
    void update_sigma_device(const std::vector<double>& sigma);
    // IForwardSolver::update_sigma (forward_solver_fv.cpp:53-58): sigma of
    // every cell from the model, then the device arrays as set_sigma.
    void update_sigma(const model::ConductivityModel& model) override;

    void compute_responses(const data::MTData& observed, data::MTData& predicted) override;
    void compute_single_frequency(int freq_idx, const data::MTData& observed,
                                  data::MTData& predicted) override;
    void solve_frequency(double freq_hz, int freq_idx, const data::MTData& observed,
                         data::MTData& predicted);
    void factorize_frequency(double freq_hz) override;
    void release_factorization() override;

    void build_adjoint_rhs_from_residual(int freq_idx,
                                         const std::vector<std::array<cplx, 4>>& weighted_residual,
                                         std::vector<cplx>& adj_rhs_pol1,
                                         std::vector<cplx>& adj_rhs_pol2) override;
    void adjoint_solve_complex(const std::vector<cplx>& rhs, std::vector<cplx>& lambda) override;
    void solve_rhs_complex(const std::vector<cplx>& rhs, std::vector<cplx>& solution) override;
    void compute_sensitivity_complex(const std::vector<cplx>& e_bg, const std::vector<cplx>& lambda,
                                     std::vector<double>& sensitivity) override;
    void build_perturbation_rhs_complex(int polarization, const std::vector<double>& delta_sigma,
                                        std::vector<cplx>& pert_rhs) override;
    void extract_delta_impedance_complex(const std::vector<cplx>& de, int polarization,
                                         std::vector<std::array<cplx, 4>>& delta_z) override;
    const std::vector<cplx>& background_E(int pol) const override { return pol == 0 ? e1_ : e2_; }
    void set_background_fields_complex(const std::vector<cplx>& e1, const std::vector<cplx>& e2) override {
        e1_ = e1;
        e2_ = e2;
    }
    void set_solver_tolerance_override(double tol) override { tol_override_ = tol; }
    double current_omega() const override { return omega_; }
    int num_elements() const override { return prob_.num_cells(); }
    // Called once per frequency as (index, count, 0, "Forward") before the
    // batch that solves them, as ForwardSolverFV forwards it
    // (forward_solver_fv.h:139-143, forward_solver_fv.cpp:73-75).
    void set_freq_progress_callback(FreqProgressCB cb) override { freq_cb_ = std::move(cb); }

    // Batched forward: both polarisations of every frequency in one device
    // call.  e_out[2*f + pol] receives the total field, z_out[f] the station
    // impedances (Zxx, Zxy, Zyx, Zyy per station).
    void solve_frequencies(const std::vector<double>& freqs_hz,
                           std::vector<std::vector<cplx>>& e_out,
                           std::vector<std::vector<std::array<cplx, 4>>>& z_out,
                           std::vector<m3d_solve_result>* results = nullptr);
    // Batched adjoint / perturbation solves: rhs[k] at omega[k].
    void adjoint_solve_batch(const std::vector<double>& omegas,
                             const std::vector<std::vector<cplx>>& rhs,
                             std::vector<std::vector<cplx>>& lambda,
                             std::vector<m3d_solve_result>* results = nullptr);

    FVProblem& problem() { return prob_; }
    const FVProblem& problem() const { return prob_; }
    void find_stations(const std::vector<std::array<double, 2>>& xy) {
        prob_.find_stations(xy);
        stations_mapped_ = true;
    }
    const std::vector<StationHCache>& station_cache() const { return cache_; }
    void set_station_cache(const std::vector<StationHCache>& c) { cache_ = c; }
    const std::vector<m3d_solve_result>& last_results() const { return last_; }
    m3d_operator* device_operator() const { return op_; }
    m3d_solve_params solve_params() const;

private:
    FVProblem prob_;
    m3d_context* ctx_ = nullptr;
    m3d_operator* op_ = nullptr;
    double omega_ = 0;
    double tol_override_ = 0;
    bool stations_mapped_ = false;
    std::vector<cplx> e1_, e2_;
    std::vector<StationHCache> cache_;
    std::vector<m3d_solve_result> last_;
    FreqProgressCB freq_cb_;

    void map_stations(const data::MTData& observed);
    void check(int status) const;
};

}  // namespace m3d
