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
constexpr int kMaxBlocks = 2048;   // grid cap for reduction kernels (8 blocks per CU)
constexpr int kMaxSys = 32;        // systems per batched launch
constexpr int kXcds = 8;           // accelerator complex dies (one L2 each)
constexpr int kTileWaves = 16;     // wavefronts of a tiled-product workgroup (one slice each)

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
    double div_norm, div_norm0, pcg_rnorm, pcg_tol;
    double2 rz, pcg_alpha, pcg_beta;
    int dc_status, pcg_status, pcg_iter, pcg_max;
    // True residual / refinement
    double true_res, true_bnorm;
    // Least-squares scale of an initial guess
    double2 guess_scale;
    int pad0, pad1;
};

// Launch shape of the row kernels at one batch width (chosen per operator,
// precision and width; see m3d_operator_autotune).  Only the assignment of
// rows to threads and workgroups changes, never the arithmetic of a row.
struct LaunchShape {
    int sweep_sub;  // lanes per row in the SSOR sweeps (1, 2, 4 or 8; width * sub <= 32)
    int sweep_cap;  // grid cap of a sweep launch (workgroups, a multiple of kXcds)
    int row_cap;    // grid cap of the row kernels that reduce (a multiple of kXcds, <= kMaxBlocks)
    int sweep_chunk;  // 1: sweeps grid-stride over row blocks; k > 1: each workgroup
                      // takes k consecutive row blocks (neighbouring rows on one CU)
    int packed;       // 1: gather from the packed blocks when the operator has them
    int wave;         // 1 (widths 8-32): couplings loaded once per row and passed among its lanes
    int tiled;        // 1 (widths 8-32): the plain product runs over the LDS-staged row tiles
};
constexpr int kWidths = 6;  // batch widths 1, 2, 4, 8, 16, 32
__host__ __device__ constexpr int width_index(int ns) {
    return ns <= 1 ? 0 : ns <= 2 ? 1 : ns <= 4 ? 2 : ns <= 8 ? 3 : ns <= 16 ? 4 : 5;
}

// Row tiles of the LDS-staged products at one batch width.  A tile is a run of
// rows consecutive in the caller's edge order (the Morton order of the mesh
// cells), so its rows gather from a compact neighbourhood: the distinct rows
// its couplings and diagonals read (stage_rows[stage_off[t] ..
// stage_off[t + 1]), ascending) are copied into LDS once, and every coupling
// then reads x from LDS by its local index.  The tile's rows are cut into
// slices of H = 64 / width rows, one wavefront each, whose entries are stored
// column-major (words[slc.x + k * H + l], k < slc.y): the local index in the
// low lbits bits and the index of the value in the operator's value table
// above them, word 0 for padding (value 0); each row keeps the entry order of
// the SELL blocks (lower couplings, then upper).  Every slice has H slots,
// slot[c * H + l] (= slot[slc.z + l]) = {device row, local index of that
// row} or {-1, 0} for a padding slot, so a slice's rows load without waiting
// for its descriptor.
struct DevTiles {
    int ntiles = 0;     // 0: no tiles at this width
    int lbits = 0;
    int max_stage = 0;  // most rows any tile stages
    const int* stage_off = nullptr;
    const int* stage_rows = nullptr;
    const int* slice_off = nullptr;
    const int4* slc = nullptr;
    const int2* slot = nullptr;
    const unsigned* words = nullptr;
};

// Device view of one mesh operator (device row order, padded to slices).
template <typename V>
struct DevOp {
    int n;                 // padded rows
    LaunchShape shape[kWidths];
    const int4* slc;       // per slice: {lower offset, lower width, upper offset, upper width}
    // Packed blocks (vtab != null): one word per entry, column in the low
    // cbits bits and the index of its value in vtab above them; vtab[0] = 0
    // for padding.  Used at the widths whose shape asks for them; the
    // separate value and column blocks otherwise.
    const unsigned* l_pk;
    const unsigned* u_pk;
    const V* vtab;
    int cbits;
    int ntab;              // values in vtab
    const V* l_val;
    const int* l_col;
    const V* u_val;
    const int* u_col;
    const V* diag_s;       // per row; 1 on constrained / padding rows
    const V* me;           // per row; 0 on constrained / padding rows
    // Per batch width: the row blocks (kBlock / width rows each) sorted by the
    // smallest caller edge index among their rows, or null.  Used by the
    // kernels without reductions, which may visit row blocks in any order.
    const int* order[kWidths];
    // Per batch width: row tiles of the LDS-staged plain product (ntiles = 0:
    // none, the product gathers from the SELL blocks).
    DevTiles tiles[kWidths];
};

// Row-block schedule: a grid stride over the row blocks, whose workgroups
// take the row blocks of a stride in windows: each full group of 64
// workgroups takes 64 consecutive row blocks, eight consecutive ones per XCD
// (the dispatcher deals workgroup i to XCD i % 8), so neighbouring rows
// share an XCD's L2 while all XCDs work on one moving window of rows, which
// the memory-side Infinity Cache serves.  window_block() is the workgroup of
// the plain grid stride whose row blocks this one takes (a bijection on the
// grid); a workgroup's block partials go to that slot, so every partial holds
// the rows it would hold in the plain grid stride, added in the same order,
// and no sum changes.  Measured on a 1.45 M-edge mesh against the plain grid
// stride: the width-2 backward sweep 10.46 -> 10.00 us per launch and a
// BiCGStab iteration 1-3.5 % shorter at every width.  A schedule giving each
// XCD a contiguous eighth of the rows (one L2 per region) measured slower on
// every solver kernel (a BiCGStab iteration 0.65 vs 0.56 ms at one system).
struct RowRange {
    int first, end, step;
};
__device__ inline int window_block() {
    const int b = blockIdx.x, g64 = int(gridDim.x) / 64 * 64;
    return b < g64 ? b / 64 * 64 + b % kXcds * 8 + b / kXcds % 8 : b;
}
__device__ inline RowRange row_range(int nrb) {
    RowRange r;
    r.first = window_block();
    r.end = nrb;
    r.step = gridDim.x;
    return r;
}

// Positions in a spatial block order for a kernel without reductions: the
// order is cut into kXcds contiguous parts and the workgroups dispatched to
// one XCD (workgroup i goes to XCD i % kXcds) walk one part, so rows that
// gather the same neighbours run on one XCD close in time and share its L2.
// The grid must be a multiple of kXcds.
__device__ inline RowRange xcd_part(int nrb) {
    const int x = blockIdx.x % kXcds;
    RowRange r;
    r.first = int(static_cast<long long>(nrb) * x / kXcds) + blockIdx.x / kXcds;
    r.end = int(static_cast<long long>(nrb) * (x + 1) / kXcds);
    r.step = gridDim.x / kXcds;
    return r;
}

// Row blocks of a sweep workgroup: the windowed grid stride, or `chunk`
// consecutive row blocks per workgroup (the grid then holds
// ceil(nrb / chunk) workgroups).
__device__ inline RowRange row_range_chunk(int nrb, int chunk) {
    if (chunk <= 1) return row_range(nrb);
    RowRange r;
    r.first = blockIdx.x * chunk;
    r.end = min(nrb, r.first + chunk);
    r.step = 1;
    return r;
}

// Bit s set when system s has status st[s].status == want.  The addresses are
// uniform, so the loads go through the scalar cache and a kernel can branch on
// a system's status without first waiting for a vector load; the status is
// written by an earlier kernel (a kernel boundary orders it).
template <int NS>
__device__ inline unsigned status_mask(const SysState* __restrict__ st, int want) {
    unsigned m = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) m |= (st[s].status == want ? 1u : 0u) << s;
    return m;
}

// Reductions run in two kernels.  The row kernel leaves one partial sum per
// (block, system, quantity); a one-block finishing kernel, ordered after it on
// the stream, adds them up and updates the solver scalars.  The kernel
// boundary makes the partials visible, so no block issues a device-scope
// fence (on CDNA that writes back the XCD's L2) or a same-address atomic.

constexpr int kFinBlock = 1024;  // threads of a finishing kernel
// Offset of the ||s||^2 partials in a workspace's partials buffer (after
// three quantities x kMaxSys systems x kMaxBlocks blocks).
constexpr size_t kPartSNorm = size_t(3) * kMaxSys * kMaxBlocks;

// Block stage: sums NQ per-thread values over the threads of each system in
// the block and stores them block-major, part[window_block() * P + p] with
// P = NQ * NS pairs, p = q * NS + sys (the slot of the rows row_range gave
// the block).  All threads must call it.
template <int NS, int NQ>
__device__ inline void block_partials(const double (&v)[NQ], double* part) {
    constexpr int P = NQ * NS;
    __shared__ double sm[kBlock / 64][P];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
        for (int off = 32; off >= NS; off >>= 1) x += __shfl_xor(x, off);
        if (lane < NS) sm[wave][q * NS + lane] = x;
    }
    __syncthreads();
    if (tid < P) {
        double acc = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) acc += sm[w][tid];
        part[size_t(window_block()) * P + tid] = acc;
    }
}

// Finish stage, in a kernel of one block of kFinBlock threads: the totals of
// the nb block partials.  Returns true for threads tid < NS, which hold the
// totals of system tid in out[q].  Each pair is summed in a fixed order, so
// results are run-to-run identical.  All threads must call it.
template <int NS, int NQ>
__device__ inline bool finish_partials(const double* __restrict__ part, int nb, double (&out)[NQ]) {
    constexpr int P = NQ * NS;
    constexpr int G = kFinBlock / P;  // thread groups per pair
    constexpr int S = G * P;          // threads that load
    __shared__ double fin[kFinBlock];
    const int tid = threadIdx.x;
    double acc = 0.0;
    if (tid < S) {
        // Thread tid owns pair tid % P and blocks tid / P, tid / P + G, ...
        // 16 loads in flight, then 4, then 1; added in increasing block order
        // either way.
        const size_t total = size_t(nb) * P;
        size_t i = tid;
        for (; i + 15 * size_t(S) < total; i += 16 * size_t(S)) {
            double a[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) a[u] = part[i + u * size_t(S)];
#pragma unroll
            for (int u = 0; u < 16; ++u) acc += a[u];
        }
        for (; i + 3 * size_t(S) < total; i += 4 * size_t(S)) {
            const double a0 = part[i], a1 = part[i + S], a2 = part[i + 2 * S], a3 = part[i + 3 * S];
            acc += a0;
            acc += a1;
            acc += a2;
            acc += a3;
        }
        for (; i < total; i += S) acc += part[i];
    }
    fin[tid] = acc;
    __syncthreads();
    // Pairwise tree over the G groups of each pair (fin[g * P + p]): group g
    // takes group g + span, for span = G2/2, G2/4, ..., 1 (G2 = G rounded up
    // to a power of two).
    constexpr int G2 = G <= 1 ? 1 : 1 << (32 - __builtin_clz(G - 1));
#pragma unroll 1
    for (int span = G2 / 2; span > 0; span >>= 1) {
        for (int idx = tid; idx < span * P; idx += kFinBlock) {
            const int g = idx / P;
            if (g + span < G) fin[idx] += fin[idx + span * P];
        }
        __syncthreads();
    }
    if (tid >= NS) return false;
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = fin[q * NS + tid];
    return true;
}

}  // namespace dev
}  // namespace m3d
