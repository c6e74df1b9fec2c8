// Shared device-side definitions: complex helpers, per-system solver state,
// the device operator view and block reductions with a deterministic
// last-block finish.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace m3d {
namespace dev {

constexpr int kBlock = 256;        // threads per block (4 wavefronts of 64)
constexpr int kSlice = 64;         // SELL slice height = one wavefront of rows
constexpr int kMaxBlocks = 1024;   // grid cap for reduction kernels (4 blocks per CU)
constexpr int kMaxSys = 32;        // systems per batched launch

// Complex value types per precision.
template <typename V> struct CT;
template <> struct CT<double> { using t = double2; };
template <> struct CT<float> { using t = float2; };
template <typename V> using cvec = typename CT<V>::t;

__host__ __device__ inline double2 mk(double a, double b) { return make_double2(a, b); }
__device__ inline double2 cmul(double2 a, double2 b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconjmul(double2 a, double2 b) {  // conj(a) * b
    return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline double cabs(double2 a) { return hypot(a.x, a.y); }
// (a + ib) / (c + id) with Smith's ordering of operations.
__device__ inline double2 cdiv(double2 num, double2 den) {
    const double a = num.x, b = num.y, c = den.x, d = den.y;
    if (fabs(c) < fabs(d)) {
        const double r = c / d, dn = (c * r) + d;
        return mk(((a * r) + b) / dn, ((b * r) - a) / dn);
    }
    const double r = d / c, dn = (d * r) + c;
    return mk(((b * r) + a) / dn, (b - (a * r)) / dn);
}
// 1 / (a + ib) for the SSOR diagonal, same operation order as cdiv(1, d).
template <typename V>
__device__ inline void cinv(V a, V b, V& ra, V& rb) {
    if (fabs(a) < fabs(b)) {
        const V r = a / b, dn = (a * r) + b;
        ra = r / dn;
        rb = V(-1) / dn;
    } else {
        const V r = b / a, dn = (b * r) + a;
        ra = V(1) / dn;
        rb = -r / dn;
    }
}

// Per-system BiCGStab / DivCorr state (double scalars in every precision).
enum Status : int {
    kRunning = 0,
    kConverged = 1,
    kConvergedEarly = 2,
    kBreakdown = 3,
    kMaxIter = 4,
    kIdle = 5,          // not part of the current phase
    kPendConvEarly = 6, // ||s|| converged: x += alpha p_hat pending
    kPendBreakTT = 7,   // (t,t) breakdown: x += alpha p_hat pending
};

struct SysState {
    double2 rho, rho_old, alpha, omega, beta;
    double bnorm, rnorm, snorm, residual;
    double tol;
    int status, iter, maxiter, info;
    // DivCorr
    double div_norm, div_norm0, pcg_rnorm;
    double2 rz, pcg_alpha, pcg_beta;
    int dc_status, pcg_status, pcg_iter, pcg_max;
    // True residual / refinement
    double true_res, true_bnorm;
    int pad0, pad1;
};

// Device view of one mesh operator (device row order, padded to slices).
template <typename V>
struct DevOp {
    int n;                 // padded rows
    const int* l_off;      // per slice: first element of the lower part
    const int* l_w;        // per slice: lower width
    const int* u_off;
    const int* u_w;
    const V* l_val;
    const int* l_col;
    const V* u_val;
    const int* u_col;
    const V* diag_s;       // per row; 1 on constrained / padding rows
    const V* me;           // per row; 0 on constrained / padding rows
};

// Sum NQ per-thread values over the threads of each system in the block and
// then over all blocks.  Returns true in exactly one block (the last to
// finish), whose threads tid < NS then hold the grid totals in out[q].
// Partials are summed in block order, so results are run-to-run identical.
template <int NS, int NQ>
__device__ inline bool grid_reduce(double (&v)[NQ], double* part, unsigned* counter,
                                   double (&out)[NQ]) {
    __shared__ double sm[kBlock / 64][NQ][NS];
    __shared__ double fin[kBlock];
    __shared__ bool last;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
        for (int off = 32; off >= NS; off >>= 1) x += __shfl_xor(x, off);
        if (lane < NS) sm[wave][q][lane] = x;
    }
    __syncthreads();
    const int nb = gridDim.x;
    if (tid < NS * NQ) {
        const int q = tid / NS, s = tid % NS;
        double acc = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) acc += sm[w][q][s];
        part[size_t(q * NS + s) * kMaxBlocks + blockIdx.x] = acc;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        const unsigned t = atomicAdd(counter, 1u);
        last = (t == unsigned(nb - 1));
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    // P = NQ*NS pairs; G groups of threads share one pair.
    constexpr int P = NQ * NS;
    constexpr int G = (kBlock / P) > 0 ? (kBlock / P) : 1;
    double acc = 0.0;
    if (tid < P * G) {
        const int p = tid % P, g = tid / P;
        const double* src = part + size_t(p) * kMaxBlocks;
        for (int b = g; b < nb; b += G)
            acc += __hip_atomic_load(src + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    fin[tid] = acc;
    __syncthreads();
    if (tid < NS) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int p = q * NS + tid;
            double s = 0.0;
            for (int g = 0; g < G; ++g) s += fin[g * P + p];
            out[q] = s;
        }
    }
    if (tid == 0) *counter = 0u;
    return true;
}

}  // namespace dev
}  // namespace m3d
