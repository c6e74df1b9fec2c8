// The forward-solver plugin interface the inversion drivers call
// (include/maple3dmt/forward/iforward_solver.h:31-115), with the conductivity
// model type its update_sigma takes (include/maple3dmt/model/
// conductivity_model.h:43-91: cell parameters, log conductivity or log
// resistivity, sigma(i) = exp(p_i) or 1 / exp(p_i)).  ForwardSolverGPU
// implements it, so a driver written against the interface switches to the
// device solve without other changes.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <functional>
#include <string>
#include <vector>

#include "mt_data.h"

namespace m3d {
namespace model {

class ConductivityModel {
public:
    enum class Parameterisation { LOG_CONDUCTIVITY, LOG_RESISTIVITY };

    // init_3d (conductivity_model.cpp:33-48): n_elements cells at sigma_background.
    void init_3d(int n_elements, double sigma_background) {
        n_elements_ = n_elements;
        params_.assign(static_cast<size_t>(n_elements),
                       type_ == Parameterisation::LOG_CONDUCTIVITY ? std::log(sigma_background)
                                                                  : std::log(1.0 / sigma_background));
    }
    int num_elements() const { return n_elements_; }
    int num_params() const { return static_cast<int>(params_.size()); }
    const std::vector<double>& params() const { return params_; }
    std::vector<double>& params() { return params_; }
    Parameterisation param_type() const { return type_; }
    void set_parameterisation(Parameterisation t) { type_ = t; }
    // conductivity_model.cpp:55-61.
    double sigma(int i) const {
        return type_ == Parameterisation::LOG_CONDUCTIVITY ? std::exp(params_[i]) : 1.0 / std::exp(params_[i]);
    }

private:
    int n_elements_ = 0;
    std::vector<double> params_;
    Parameterisation type_ = Parameterisation::LOG_CONDUCTIVITY;
};

}  // namespace model

namespace forward {

using cplx = std::complex<double>;

// iforward_solver.h:31-115, member for member.
class IForwardSolver {
public:
    virtual ~IForwardSolver() = default;
    virtual void compute_responses(const data::MTData& observed, data::MTData& predicted) = 0;
    virtual void compute_single_frequency(int freq_idx, const data::MTData& observed,
                                          data::MTData& predicted) = 0;
    virtual void factorize_frequency(double freq_hz) = 0;
    virtual void release_factorization() = 0;
    virtual void update_sigma(const model::ConductivityModel& model) = 0;
    virtual int num_elements() const = 0;
    virtual double current_omega() const = 0;
    virtual void build_adjoint_rhs_from_residual(int freq_idx,
                                                 const std::vector<std::array<cplx, 4>>& weighted_residual,
                                                 std::vector<cplx>& adj_rhs_pol1,
                                                 std::vector<cplx>& adj_rhs_pol2) = 0;
    virtual void adjoint_solve_complex(const std::vector<cplx>& rhs, std::vector<cplx>& lambda) = 0;
    virtual void compute_sensitivity_complex(const std::vector<cplx>& e_bg, const std::vector<cplx>& lambda,
                                             std::vector<double>& sensitivity) = 0;
    virtual void build_perturbation_rhs_complex(int polarization, const std::vector<double>& delta_sigma,
                                                std::vector<cplx>& pert_rhs) = 0;
    virtual void solve_rhs_complex(const std::vector<cplx>& rhs, std::vector<cplx>& solution) = 0;
    virtual void extract_delta_impedance_complex(const std::vector<cplx>& de, int polarization,
                                                 std::vector<std::array<cplx, 4>>& delta_z) = 0;
    virtual const std::vector<cplx>& background_E(int pol) const = 0;
    virtual void set_background_fields_complex(const std::vector<cplx>& e1, const std::vector<cplx>& e2) = 0;
    // Default no-op, as in the interface (:108-110).
    virtual void set_solver_tolerance_override(double /*tol*/) {}
    using FreqProgressCB = std::function<void(int, int, double, const std::string&)>;
    virtual void set_freq_progress_callback(FreqProgressCB cb) = 0;
};

}  // namespace forward
}  // namespace m3d
