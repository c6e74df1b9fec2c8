// Host-callable launchers of the solve kernels (kernels.hip).
#pragma once

#include "dev_common.h"

namespace m3d {
namespace dev {

// Misfit gradient in ln sigma summed over nsys systems (one thread per cell).
void launch_sensitivity(hipStream_t s, int nc, int n, int nsys, const int* cptr, const int* cedge,
                        const double* wgt, const double* sigma, const double* omega, const double2* E,
                        const double2* lam, double* g);

// Input of a forward sweep: in as given, the updated search direction
// p = r + beta (p - omega v), or s = r - alpha v (both written out).
enum SweepInput : int { kInPlain = 0, kInPUpdate = 1, kInSUpdate = 2 };

template <int NS, typename V>
struct Launch {
    using C = cvec<V>;
    static void pack(hipStream_t s, int npad, int n, int nsys, const int* perm, const double2* src,
                     C* dst);
    static void unpack(hipStream_t s, int npad, int n, int nsys, const int* perm, const C* src,
                       double2* dst);
    static void bicg_init(hipStream_t s, const DevOp<V>& op, const double* omega, const C* b,
                          const C* x, C* r, C* r0, SysState* st, double* part,
                          int x_zero);
    // mode: SweepInput (in = r and p the output for the updates).  zl !=
    // nullptr (the last colour only): also write the backward result z = y
    // of these rows, which have no later colours.
    static void sweep_fwd(hipStream_t s, const DevOp<V>& op, const double* omega, int r0, int r1,
                          int mode, const C* in, C* p, const C* v, C* y, const SysState* st,
                          C* zl = nullptr);
    static void sweep_bwd(hipStream_t s, const DevOp<V>& op, const double* omega, int r0, int r1,
                          const C* y, C* z, const SysState* st);
    // ys != nullptr: z is SSOR-preconditioned and ys its forward-sweep output
    // (Eisenstat form, lower couplings only); nullptr: plain A z.
    static void spmv_v(hipStream_t s, const DevOp<V>& op, const double* omega, const C* z,
                       const C* ys, const C* r0, C* v, SysState* st, double* part);
    static void spmv_t(hipStream_t s, const DevOp<V>& op, const double* omega, const C* z,
                       const C* ys, const C* sv, C* t, SysState* st, double* part);
    static void bicg_xr(hipStream_t s, int n, C* x, const C* ph, const C* sh, const C* sv,
                        const C* t, const C* r0, C* r, SysState* st, double* part);
    static void spmv(hipStream_t s, const DevOp<V>& op, const double* omega, const C* x, const C* b,
                     C* y);
    static void fill_zero(hipStream_t s, size_t count, C* x);
    static void interp_hanging(hipStream_t s, int npairs, const int* tri, C* x);
};

template <int NS>
struct Launch64 {
    static void true_res(hipStream_t s, const DevOp<double>& op, const double* omega,
                         const double2* x, const double2* b, double2* r, SysState* st,
                         double* part);
    // x *= gamma per system, gamma = (A x, b) / (A x, A x).
    static void scale_guess(hipStream_t s, const DevOp<double>& op, const double* omega,
                            const double2* b, double2* x, SysState* st, double* part);
    static void zero_systems(hipStream_t s, int npad, const int* flags, double2* x);
    // dst = src (0 when src is null) on the flagged systems.
    static void copy_systems(hipStream_t s, int npad, const int* flags, const double2* src, double2* dst);
    static void forward_total(hipStream_t s, int npad, const unsigned char* bc,
                              const double2* epri, const double2* b, double2* x);
    static void add_correction(hipStream_t s, size_t count, const float2* d, double2* x);
    static void fix_constrained(hipStream_t s, int npad, const unsigned char* cons, const double2* b,
                                double2* x);
    static void demote(hipStream_t s, size_t count, const double2* src, const double* scale,
                       float2* dst);
    static void add_scaled(hipStream_t s, size_t count, const float2* d, const double* scale,
                           double2* x);
    static void phi0(hipStream_t s, int nn, const int* nt_ptr, const int* nt_edge,
                     const double* nt_val, const unsigned char* dead, const double* omega,
                     const double2* b, double2* phi0);
};

// The DivCorr PCG (divcorr_poisson_, forward_solver_fv.cpp:623-722) in value
// type V: double, or float for the mixed-precision path's interleaved
// corrections (complex64 node vectors, FP32 Laplacian; reductions in double).
// The field E and phi0 stay complex128 in both.
template <int NS, typename V>
struct LaunchDc {
    using C = cvec<V>;
    static void divj(hipStream_t s, int nn, int it, double tol, int cg_max, double pcg_tol, const int* nt_ptr,
                     const int* nt_edge, const double* nt_val, const double* ms, const unsigned char* bc_node,
                     const V* ldiag, const double2* E, const double2* phi0, C* r, C* z, C* p, C* phi,
                     SysState* st, double* part);
    static void lp(hipStream_t s, int nn, const int* lptr, const int* lcol, const V* lval, const C* p, C* ap,
                   SysState* st, double* part);
    static void update(hipStream_t s, int nn, const V* ldiag, const C* p, const C* ap, C* phi, C* r,
                       SysState* st, double* part);
    // p = D^-1 r + beta p (the Jacobi-preconditioned residual formed in place).
    static void dir(hipStream_t s, int nn, const V* ldiag, const C* r, C* p, const SysState* st);
    static void correct(hipStream_t s, int npad, const int* g_col, const double* g_val,
                        const unsigned char* bc_edge, const C* phi, double2* E, const SysState* st);
};

void launch_bicg_begin(hipStream_t s, SysState* st, unsigned mask, int maxiter, double tol);
// Gauss-Newton perturbation right-hand side, caller order [nsys][n].
void launch_perturbation_rhs(hipStream_t s, int n, int nsys, int nc, const int* e_ptr, const int* e_cells,
                             const int* adj_ptr, const int* adj_cell, const double* len, const double* sigma,
                             const double* dsig, const double* omega, const double2* ebg, double2* rhs);
// Sigma update: per device row the operator's Me (FP64, FP32) and the
// DivCorr weight; per node the Laplacian values and Jacobi diagonal.
void launch_sigma_edges(hipStream_t s, int npad, int nc, const int* perm, const int* adj_ptr,
                        const int* adj_cell, const double* area, const double* len,
                        const unsigned char* massless, const unsigned char* dead,
                        const unsigned char* constrained, const unsigned char* dc_dead,
                        const unsigned char* in_pattern, const double* sigma, double* me, float* me32,
                        double* ms, int* err);
void launch_sigma_laplacian(hipStream_t s, int nn, const int* nt_ptr, const int* nt_edge, const double* nt_val,
                            const double* ms, const int* lptr, const int* lcol, const int* ledge,
                            const int* g_col, const double* g_val, const unsigned char* bc_node, double* lval,
                            double* ldiag);
template <typename V>
void launch_regroup(hipStream_t s, int rows, int na, int nb, const int* map, cvec<V>* wide,
                    cvec<V>* narrow, bool to_narrow);
void launch_dc_begin(hipStream_t s, SysState* st, unsigned mask);
// dst[i] = float(src[i]).
void launch_to_f32(hipStream_t s, size_t n, const double* src, float* dst);

}  // namespace dev
}  // namespace m3d
