// HIP kernels of the batched FP64 / FP32 octree curl-curl solve on CDNA4.
//
// Layout: the operator is kept in colour order, padded so every colour starts
// on a 64-row slice; each slice stores its strictly lower (earlier colours)
// and strictly upper (later colours) couplings as column-major SELL-64 blocks,
// so lane k of a wavefront reads entry j of row k with one coalesced load.
// Vectors of a batch are system-minor, x[row * NS + sys], so the NS lanes of
// one row read NS consecutive complex values when they gather a column.
// Thread t of a block works on row (t / NS), system (t % NS).
#include <algorithm>

#include "dev_common.h"
#include "kernels.h"

namespace m3d {
namespace dev {

namespace {

template <int NS>
__device__ inline int rows_per_block() { return kBlock / NS; }

// Descriptor of the slice holding a wavefront's rows.  Every kernel maps a
// wavefront onto consecutive rows of one 64-row slice (a wavefront covers
// 64 / (NS * SUB) rows starting at a multiple of that, and colours and the
// padded row count are multiples of 64), so the index is uniform: read from
// the first lane, the load goes through the scalar cache and does not wait
// behind the wavefront's vector loads.
__device__ inline int4 slice_desc(const int4* __restrict__ slc, int row) {
    return slc[__builtin_amdgcn_readfirstlane(row) >> 6];
}

// Entries k = k0, k0 + STEP, ... < w of one row of a SELL block, added into
// (ar, ai) in increasing k.  The column indices and values of U entries are
// loaded before any of their x entries, so a row waits for two memory round
// trips per U entries instead of two per entry.  Every lane of a wavefront
// works on rows of one slice, so off and w are read from the first lane into
// scalar registers and the loop runs without divergence; entries past w load
// nothing and are added as 0 * 0 (a row sum is never -0, so adding +0 leaves
// it unchanged, as the entries' absence would).
template <int U, int STEP, typename V>
__device__ inline void gather_sell(const V* __restrict__ val, const int* __restrict__ col, int off,
                                   int w, int lane, int k0, const cvec<V>* __restrict__ x, int NS,
                                   int sys, V& ar, V& ai) {
    off = __builtin_amdgcn_readfirstlane(off);
    w = __builtin_amdgcn_readfirstlane(w);
    for (int k = k0; k < w; k += U * STEP) {
        int c[U];
        V a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k + u * STEP;
            const int idx = off + kk * kSlice + lane;
            c[u] = kk < w ? col[idx] : 0;
            a[u] = kk < w ? val[idx] : V(0);
        }
        cvec<V> xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u].x = V(0);
            xv[u].y = V(0);
            if (k + u * STEP < w) xv[u] = x[size_t(c[u]) * NS + sys];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ar += a[u] * xv[u].x;
            ai += a[u] * xv[u].y;
        }
    }
}

// gather_sell over a packed block: the column and the value's table index
// come from one word, so an entry costs 4 bytes instead of 8 + 4 (or 4 + 4).
// Entries past w load nothing and add 0 * 0, as in gather_sell.
// The additions are those of gather_sell in the same order.
template <int U, int STEP, typename V>
__device__ inline void gather_packed(const unsigned* __restrict__ pk, const V* __restrict__ tab, int cbits,
                                     int off, int w, int lane, int k0, const cvec<V>* __restrict__ x,
                                     int NS, int sys, V& ar, V& ai) {
    off = __builtin_amdgcn_readfirstlane(off);
    w = __builtin_amdgcn_readfirstlane(w);
    const unsigned cm = (1u << cbits) - 1u;
    // The words of the next U entries load together with this batch's
    // gathers (one memory round trip per batch after the first, not two).
    unsigned wd[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int kk = k0 + u * STEP;
        wd[u] = kk < w ? pk[off + kk * kSlice + lane] : 0u;
    }
    for (int k = k0; k < w; k += U * STEP) {
        cvec<V> xv[U];
        V a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u].x = V(0);
            xv[u].y = V(0);
            a[u] = V(0);
            if (k + u * STEP < w) {
                xv[u] = x[size_t(wd[u] & cm) * NS + sys];
                a[u] = tab[wd[u] >> cbits];
            }
        }
        const int kn = k + U * STEP;
        unsigned nw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = kn + u * STEP;
            nw[u] = kk < w ? pk[off + kk * kSlice + lane] : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ar += a[u] * xv[u].x;
            ai += a[u] * xv[u].y;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) wd[u] = nw[u];
    }
}

// Entries k0, k0 + STEP, ... of the lower (UP = false) or upper block of a
// row, from the packed blocks when the operator has them and the launch
// shape of this batch width asks for them (a uniform branch).
template <int U, int STEP, bool UP, typename V>
__device__ inline void gather_block(const DevOp<V>& op, int off, int w, int lane, int k0,
                                    const cvec<V>* __restrict__ x, int NS, int sys, V& ar, V& ai) {
    if (op.vtab && op.shape[width_index(NS)].packed)
        gather_packed<U, STEP, V>(UP ? op.u_pk : op.l_pk, op.vtab, op.cbits, off, w, lane, k0, x, NS, sys, ar,
                                  ai);
    else
        gather_sell<U, STEP, V>(UP ? op.u_val : op.l_val, UP ? op.u_col : op.l_col, off, w, lane, k0, x, NS,
                                sys, ar, ai);
}

template <bool UP, typename V>
__device__ inline void row_gather(const DevOp<V>& op, int off, int w, int lane, const cvec<V>* __restrict__ x,
                                  int NS, int sys, V& ar, V& ai) {
    gather_block<8, 1, UP, V>(op, off, w, lane, 0, x, NS, sys, ar, ai);
}

// Wide batches (NS >= 8): the NS lanes of a row (NS consecutive lanes of a
// wavefront, vectors being system-minor) load the couplings of NS consecutive
// entries of the row at once, lane j the coupling of entry k0 + j, and hand
// each entry's coupling to the row's lanes through ds_bpermute (the LDS
// crossbar; no memory access).  An entry then costs the wavefront one vector
// memory load, its x gather, instead of three (column or packed word, value,
// x), and every NS entries one more.  The values of a packed block come from a
// copy of the value table in LDS.  rl = the row's lane within its slice, gl =
// the row's first lane within the wavefront.  Entries are added in the order
// of gather_sell / gather_packed, increasing k (padding adds 0 * 0, which
// leaves a row sum unchanged), so the sums are the same bit for bit.
template <int NS>
__host__ __device__ constexpr bool wave_gather_width() { return NS >= 8 && NS <= 32; }

__device__ inline unsigned lane_get(unsigned v, int src) {
    return unsigned(__builtin_amdgcn_ds_bpermute(src << 2, int(v)));
}
__device__ inline int lane_get(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }
__device__ inline float lane_get(float v, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ inline double lane_get(double v, int src) {
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
    return __hiloint2double(hi, lo);
}

// One lane's share of a chunk of a SELL block: the coupling of entry k of
// its row (zero past the width).
template <typename V>
struct LaneCoupling {
    unsigned word;   // packed word, or the column of a separate block
    V val;           // value of a separate block (unused when packed)
};

template <bool PACKED, typename V>
__device__ inline LaneCoupling<V> lane_load(const unsigned* __restrict__ pk, const V* __restrict__ val,
                                            const int* __restrict__ col, int off, int w, int k, int rl) {
    LaneCoupling<V> c{0u, V(0)};
    if (k < w) {
        const int idx = off + k * kSlice + rl;
        if (PACKED) {
            c.word = pk[idx];
        } else {
            c.word = unsigned(col[idx]);
            c.val = val[idx];
        }
    }
    return c;
}

// Adds the n (<= NS, uniform) entries of one chunk, whose couplings lane
// gl + j holds for entry j, in batches of U gathers in flight.  Every lane of
// the row passes its coupling on; a lane whose system is not running (run =
// false) gathers nothing and adds zeros.
template <int U, int NS, bool PACKED, typename V>
__device__ inline void lane_chunk(const LaneCoupling<V>& mine, int n, const V* __restrict__ ltab, int cbits,
                                  int gl, const cvec<V>* __restrict__ x, int sys, bool run, V& ar, V& ai) {
    static_assert(NS % U == 0, "a batch stays within the row's lanes");
    const unsigned cm = (1u << cbits) - 1u;
    for (int j = 0; j < n; j += U) {
        unsigned wd[U];
        V a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            wd[u] = lane_get(mine.word, gl + j + u);
            if (!PACKED) a[u] = lane_get(mine.val, gl + j + u);
        }
        cvec<V> xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u].x = V(0);
            xv[u].y = V(0);
            if (PACKED) a[u] = V(0);
            if (run && j + u < n) {
                xv[u] = x[size_t(PACKED ? (wd[u] & cm) : wd[u]) * NS + sys];
                if (PACKED) a[u] = ltab[wd[u] >> cbits];
            } else if (!PACKED) {
                a[u] = V(0);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ar += a[u] * xv[u].x;
            ai += a[u] * xv[u].y;
        }
    }
}

// One block of a row (UP: the upper block), chunk by chunk.
template <int NS, bool PACKED, bool UP, typename V>
__device__ inline void lane_block_at(const DevOp<V>& op, const V* __restrict__ ltab, int off, int w, int rl,
                                     int gl, const cvec<V>* __restrict__ x, int sys, bool run, V& ar, V& ai) {
    off = __builtin_amdgcn_readfirstlane(off);
    w = __builtin_amdgcn_readfirstlane(w);
    const unsigned* pk = UP ? op.u_pk : op.l_pk;
    const V* val = UP ? op.u_val : op.l_val;
    const int* col = UP ? op.u_col : op.l_col;
    if constexpr (NS <= 8) {
        // Blocks span several chunks of 8: the next chunk's couplings load
        // before this chunk's gathers (a width-8 iteration 1.3-1.7 % shorter;
        // at 16 systems, where most blocks fit one chunk, the extra load costs
        // more than it hides).
        LaneCoupling<V> c = lane_load<PACKED, V>(pk, val, col, off, w, sys, rl);
        for (int k0 = 0; k0 < w; k0 += NS) {
            const LaneCoupling<V> cn = lane_load<PACKED, V>(pk, val, col, off, w, k0 + NS + sys, rl);
            lane_chunk<8, NS, PACKED, V>(c, min(NS, w - k0), ltab, op.cbits, gl, x, sys, run, ar, ai);
            c = cn;
        }
    } else {
        for (int k0 = 0; k0 < w; k0 += NS) {
            const LaneCoupling<V> c = lane_load<PACKED, V>(pk, val, col, off, w, k0 + sys, rl);
            lane_chunk<8, NS, PACKED, V>(c, min(NS, w - k0), ltab, op.cbits, gl, x, sys, run, ar, ai);
        }
    }
}

// lane_block_at from the packed blocks when ltab (the LDS value table) is
// given, else from the separate blocks.
template <int NS, bool UP, typename V>
__device__ inline void lane_block(const DevOp<V>& op, const V* __restrict__ ltab, int off, int w, int row,
                                  const cvec<V>* __restrict__ x, int sys, bool run, V& ar, V& ai) {
    const int gl = (threadIdx.x & 63) - sys;
    if (ltab)
        lane_block_at<NS, true, UP, V>(op, ltab, off, w, row & 63, gl, x, sys, run, ar, ai);
    else
        lane_block_at<NS, false, UP, V>(op, ltab, off, w, row & 63, gl, x, sys, run, ar, ai);
}

// Lower then upper block of a row; the first chunk of both blocks is loaded
// before either is gathered, so a row with both widths <= NS waits for two
// memory round trips (couplings, then x) instead of four.
template <int NS, bool PACKED, typename V>
__device__ inline void lane_row(const DevOp<V>& op, const V* __restrict__ ltab, int4 sc, int rl, int gl,
                                const cvec<V>* __restrict__ x, int sys, bool run, V& ar, V& ai) {
    constexpr int U = 8;
    const int lo = __builtin_amdgcn_readfirstlane(sc.x), lw = __builtin_amdgcn_readfirstlane(sc.y);
    const int uo = __builtin_amdgcn_readfirstlane(sc.z), uw = __builtin_amdgcn_readfirstlane(sc.w);
    const LaneCoupling<V> l0 = lane_load<PACKED, V>(op.l_pk, op.l_val, op.l_col, lo, lw, sys, rl);
    const LaneCoupling<V> u0 = lane_load<PACKED, V>(op.u_pk, op.u_val, op.u_col, uo, uw, sys, rl);
    lane_chunk<U, NS, PACKED, V>(l0, min(NS, lw), ltab, op.cbits, gl, x, sys, run, ar, ai);
    for (int k0 = NS; k0 < lw; k0 += NS) {
        const LaneCoupling<V> c = lane_load<PACKED, V>(op.l_pk, op.l_val, op.l_col, lo, lw, k0 + sys, rl);
        lane_chunk<U, NS, PACKED, V>(c, min(NS, lw - k0), ltab, op.cbits, gl, x, sys, run, ar, ai);
    }
    lane_chunk<U, NS, PACKED, V>(u0, min(NS, uw), ltab, op.cbits, gl, x, sys, run, ar, ai);
    for (int k0 = NS; k0 < uw; k0 += NS) {
        const LaneCoupling<V> c = lane_load<PACKED, V>(op.u_pk, op.u_val, op.u_col, uo, uw, k0 + sys, rl);
        lane_chunk<U, NS, PACKED, V>(c, min(NS, uw - k0), ltab, op.cbits, gl, x, sys, run, ar, ai);
    }
}

// (A x)_row of apply_row by couplings shared among the row's lanes; ltab =
// the LDS copy of the value table when this width gathers from the packed
// blocks.
template <int NS, typename V>
__device__ inline cvec<V> apply_row_wave(const DevOp<V>& op, const V* __restrict__ ltab, int row, int sys, V w,
                                         const cvec<V>* __restrict__ x, bool run = true) {
    const int4 sc = slice_desc(op.slc, row);
    const int rl = row & 63;
    const int gl = (threadIdx.x & 63) - sys;
    V ar = 0, ai = 0;
    if (ltab)
        lane_row<NS, true, V>(op, ltab, sc, rl, gl, x, sys, run, ar, ai);
    else
        lane_row<NS, false, V>(op, ltab, sc, rl, gl, x, sys, run, ar, ai);
    const V dr = op.diag_s[row], di = w * op.me[row];
    const cvec<V> xi = x[size_t(row) * NS + sys];
    cvec<V> y;
    y.x = ar + (dr * xi.x - di * xi.y);
    y.y = ai + (dr * xi.y + di * xi.x);
    return y;
}

// The dynamic LDS of a launch (one declaration shared by every kernel).
__device__ inline unsigned char* dyn_lds() {
    extern __shared__ __align__(16) unsigned char m3d_dyn_lds[];
    return m3d_dyn_lds;
}

// The value table of a packed operator copied into LDS by the whole block
// (all threads call it); null when this width gathers from the separate
// blocks.  The launch gives the block ntab * sizeof(V) bytes of dynamic LDS.
template <int NS, typename V>
__device__ inline const V* stage_value_table(const DevOp<V>& op, unsigned char* lds) {
    if (!(op.vtab && op.shape[width_index(NS)].packed)) return nullptr;
    V* t = reinterpret_cast<V*>(lds);
    for (int i = threadIdx.x; i < op.ntab; i += blockDim.x) t[i] = op.vtab[i];
    __syncthreads();
    return t;
}

// Sweeps may split each row over SUB lanes (entries k = sub, sub + SUB, ...;
// lanes of one row are NS apart), so that one colour of ~1e5 rows still puts
// enough wavefronts on the GPU to hide gather latency at narrow batch widths.
// SUB is part of the operator's launch shape.
template <int NS, int SUB>
__host__ __device__ constexpr int sweep_rpb() { return kBlock / (NS * SUB); }

template <int SUB, bool UP, typename V>
__device__ inline void row_gather_split(const DevOp<V>& op, int off, int w, int lane,
                                        const cvec<V>* __restrict__ x, int NS, int sys, int sub, V& ar,
                                        V& ai) {
    gather_block<(SUB == 1 ? 8 : SUB >= 8 ? 2 : 4), SUB, UP, V>(op, off, w, lane, sub, x, NS, sys, ar, ai);
#pragma unroll
    for (int m = 1; m < SUB; m <<= 1) {
        ar += __shfl_xor(ar, NS * m);
        ai += __shfl_xor(ai, NS * m);
    }
}

// (A x)_row for one system: lower + upper couplings plus the diagonal
// (diag_s + i*omega*me) * x_row.
template <typename V>
__device__ inline cvec<V> apply_row(const DevOp<V>& op, int row, int NS, int sys, V w,
                                    const cvec<V>* __restrict__ x) {
    const int4 sc = slice_desc(op.slc, row);
    const int lane = row & 63;
    V ar = 0, ai = 0;
    row_gather<false>(op, sc.x, sc.y, lane, x, NS, sys, ar, ai);
    row_gather<true>(op, sc.z, sc.w, lane, x, NS, sys, ar, ai);
    const V dr = op.diag_s[row], di = w * op.me[row];
    const cvec<V> xi = x[size_t(row) * NS + sys];
    cvec<V> y;
    y.x = ar + (dr * xi.x - di * xi.y);
    y.y = ai + (dr * xi.y + di * xi.x);
    return y;
}

// (A z)_row for the SSOR-preconditioned direction z = (D+U)^-1 D y with
// y = (D+L)^-1 p, by Eisenstat's identity A z = L z + (D+U) z = L z + D y:
// only the lower (earlier-colour) couplings are gathered.
template <typename V>
__device__ inline cvec<V> apply_row_eis(const DevOp<V>& op, int row, int NS, int sys, V w,
                                        const cvec<V>* __restrict__ z, const cvec<V>* __restrict__ y) {
    const int4 sc = slice_desc(op.slc, row);
    const int lane = row & 63;
    V ar = 0, ai = 0;
    row_gather<false>(op, sc.x, sc.y, lane, z, NS, sys, ar, ai);
    const V dr = op.diag_s[row], di = w * op.me[row];
    const cvec<V> yi = y[size_t(row) * NS + sys];
    cvec<V> out;
    out.x = ar + (dr * yi.x - di * yi.y);
    out.y = ai + (dr * yi.y + di * yi.x);
    return out;
}

// apply_row_eis by couplings shared among the row's lanes (wide batches).
template <int NS, typename V>
__device__ inline cvec<V> apply_row_eis_wave(const DevOp<V>& op, const V* __restrict__ ltab, int row, int sys,
                                             V w, const cvec<V>* __restrict__ z, const cvec<V>* __restrict__ y,
                                             bool run) {
    const int4 sc = slice_desc(op.slc, row);
    V ar = 0, ai = 0;
    lane_block<NS, false, V>(op, ltab, sc.x, sc.y, row, z, sys, run, ar, ai);
    const V dr = op.diag_s[row], di = w * op.me[row];
    const cvec<V> yi = y[size_t(row) * NS + sys];
    cvec<V> out;
    out.x = ar + (dr * yi.x - di * yi.y);
    out.y = ai + (dr * yi.y + di * yi.x);
    return out;
}

__device__ inline double fmav(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fmav(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename V>
__device__ inline cvec<V> to_v(double2 a) {
    cvec<V> r;
    r.x = V(a.x);
    r.y = V(a.y);
    return r;
}
template <typename V>
__device__ inline double2 to_d(cvec<V> a) {
    return mk(double(a.x), double(a.y));
}

}  // namespace

// ---------------------------------------------------------------------------
// Layout conversion: system-major double (caller order) <-> system-minor V
// (device row order).  perm[row] = caller edge of device row, -1 = padding.
// ---------------------------------------------------------------------------
template <int NS, typename V>
__global__ void k_pack(int npad, int n, int nsys, const int* __restrict__ perm,
                       const double2* __restrict__ src, cvec<V>* __restrict__ dst) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int row = int(t / NS), sys = int(t % NS);
        const int e = perm[row];
        double2 v = mk(0.0, 0.0);
        if (e >= 0 && sys < nsys && src) v = src[size_t(sys) * n + e];
        dst[t] = to_v<V>(v);
    }
}

template <int NS, typename V>
__global__ void k_unpack(int npad, int n, int nsys, const int* __restrict__ perm,
                         const cvec<V>* __restrict__ src, double2* __restrict__ dst) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int row = int(t / NS), sys = int(t % NS);
        const int e = perm[row];
        if (e >= 0 && sys < nsys) dst[size_t(sys) * n + e] = to_d<V>(src[t]);
    }
}

// ---------------------------------------------------------------------------
// BiCGStab (bicgstab.cpp:15-166), right preconditioned, Hermitian dots.
// ---------------------------------------------------------------------------

// Start a solve on the systems whose mask bit is set.
__global__ void k_bicg_begin(SysState* st, unsigned mask, int maxiter, double tol) {
    const int s = threadIdx.x;
    if (s >= kMaxSys) return;
    SysState& S = st[s];
    if (mask & (1u << s)) {
        S.status = kRunning;
        S.iter = 0;
        S.maxiter = maxiter;
        S.tol = tol;
        S.info = 0;
        S.rho_old = mk(1.0, 0.0);
        S.alpha = mk(1.0, 0.0);
        S.omega = mk(1.0, 0.0);
        S.beta = mk(0.0, 0.0);
    } else if (S.status == kRunning || S.status == kPendConvEarly || S.status == kPendBreakTT) {
        S.status = kIdle;
    }
}

// r = b - A x (or r = b when x is known to be zero), r0 = r; norms and
// rho = (r0, r) for the first iteration (bicgstab.cpp:29-46, 51-58).
template <int NS, typename V>
__global__ void k_bicg_init(DevOp<V> op, const double* __restrict__ omega,
                            const cvec<V>* __restrict__ b, const cvec<V>* __restrict__ x,
                            cvec<V>* __restrict__ r, cvec<V>* __restrict__ r0,
                            const SysState* __restrict__ st, double* part, int x_zero) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const bool run = (status_mask<NS>(st, kRunning) >> sys) & 1u;
    const V w = V(omega[sys]);
    double acc[3] = {0.0, 0.0, 0.0};
    const int nrb = (op.n + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = rb * rpb + tid / NS;
        if (row >= op.n || !run) continue;
        const size_t i = size_t(row) * NS + sys;
        const cvec<V> bi = b[i];
        cvec<V> ri = bi;
        if (!x_zero) {
            const cvec<V> ax = apply_row(op, row, NS, sys, w, x);
            ri.x = bi.x - ax.x;
            ri.y = bi.y - ax.y;
        }
        r[i] = ri;
        r0[i] = ri;
        const double2 bd = to_d<V>(bi), rd = to_d<V>(ri);
        acc[0] += cabs2(bd);
        acc[1] += cabs2(rd);
        const double2 rr = cconjmul(rd, rd);
        acc[2] += rr.x;
    }
    block_partials<NS, 3>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_bicg_init_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[3];
    if (!finish_partials<NS, 3>(part, nb, out)) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.status != kRunning) return;
    double bn = sqrt(out[0]);
    if (bn < 1e-30) bn = 1.0;
    S.bnorm = bn;
    S.rnorm = sqrt(out[1]);
    if (S.rnorm / bn < S.tol) {
        S.status = kConverged;
        S.residual = S.rnorm / bn;
        S.iter = 0;
        S.info = 0;
        return;
    }
    S.rho = mk(out[2], 0.0);
    if (cabs(S.rho) < 1e-30) {
        S.status = kBreakdown;
        S.info = 2;
        S.iter = 1;
        S.residual = S.rnorm / bn;
    }
}

// Forward multicolour Gauss-Seidel sweep over one colour [r0, r1):
//   y_i = (in_i - sum_{j in earlier colours} a_ij y_j) / d_i   (ssor.h:58-65)
// The input, by MODE:
//   kInPlain   in
//   kInPUpdate the search direction, updated first (in = r, out p):
//              p = r (first iteration) or p = r + beta (p - omega v)
//              (bicgstab.cpp:60-67)
//   kInSUpdate s = r - alpha v (in = r, out s; bicgstab.cpp:88-92), which
//              the sweep needs row by row, so it is formed here instead of
//              in a pass of its own; its norm is reduced by k_spmv_t.
// With zl (the last colour, which has no later colours) the backward step of
// these rows is done here too: z_i = y_i - (0 / d_i), the backward sweep's
// arithmetic with an empty upper block.
// LANES (wide batches, one lane per row): couplings shared among the row's
// lanes; the lanes of systems that are not running stay to pass their
// couplings on and store nothing.
template <int NS, typename V, int SUB, int MODE, bool LANES>
__global__ void k_sweep_fwd(DevOp<V> op, const double* __restrict__ omega, int r0, int r1,
                            const cvec<V>* __restrict__ in, cvec<V>* __restrict__ p,
                            const cvec<V>* __restrict__ v, cvec<V>* __restrict__ y,
                            const SysState* __restrict__ st, cvec<V>* __restrict__ zl) {
    static_assert(!LANES || SUB == 1, "shared couplings at one lane per row");
    const int tid = threadIdx.x, sys = tid % NS, sub = (tid / NS) % SUB;
    constexpr int rpb = sweep_rpb<NS, SUB>();
    const V* ltab = LANES ? stage_value_table<NS, V>(op, dyn_lds()) : nullptr;
    const bool run = (status_mask<NS>(st, kRunning) >> sys) & 1u;
    if (!LANES && !run) return;
    unsigned first_m = 0;
#pragma unroll
    for (int q = 0; q < NS; ++q) first_m |= (st[q].iter == 0 ? 1u : 0u) << q;
    const bool first = (first_m >> sys) & 1u;
    const SysState& S = st[sys];
    const V w = V(omega[sys]);
    const int nrows = r1 - r0;
    const int nrb = (nrows + rpb - 1) / rpb;
    const V br = V(S.beta.x), bi = V(S.beta.y), wr = V(S.omega.x), wi = V(S.omega.y);
    const RowRange R = row_range_chunk(nrb, op.shape[width_index(NS)].sweep_chunk);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = r0 + rb * rpb + tid / (NS * SUB);
        if (row >= r1) continue;
        const size_t i = size_t(row) * NS + sys;
        // Every load the row needs besides the gather goes out before it.
        const int4 sc = slice_desc(op.slc, row);
        const V ds = op.diag_s[row], me = op.me[row];
        cvec<V> src = in[i];
        if (MODE == kInPUpdate) {
            if (!first) {
                const cvec<V> pi = p[i], vi = v[i];
                // p - omega v
                const V tr = pi.x - (wr * vi.x - wi * vi.y);
                const V ti = pi.y - (wr * vi.y + wi * vi.x);
                src.x = src.x + (br * tr - bi * ti);
                src.y = src.y + (br * ti + bi * tr);
            }
            if (sub == 0 && run) p[i] = src;
        } else if (MODE == kInSUpdate) {
            // alpha v with explicit fused multiply-adds: the rounding of s
            // is fixed by the source, not by what contraction the compiler
            // picks for this kernel.
            const V ar = V(S.alpha.x), ai = V(S.alpha.y);
            const cvec<V> vi = v[i];
            const V avr = fmav(ar, vi.x, -(ai * vi.y));
            const V avi = fmav(ai, vi.x, ar * vi.y);
            src.x = src.x - avr;
            src.y = src.y - avi;
            if (sub == 0 && run) p[i] = src;
        }
        V ar = 0, ai = 0;
        if constexpr (LANES)
            lane_block<NS, false, V>(op, ltab, sc.x, sc.y, row, y, sys, run, ar, ai);
        else
            row_gather_split<SUB, false>(op, sc.x, sc.y, row & 63, y, NS, sys, sub, ar, ai);
        V dr, di;
        cinv(ds, w * me, dr, di);
        const V er = src.x - ar, ei = src.y - ai;
        cvec<V> out;
        out.x = er * dr - ei * di;
        out.y = er * di + ei * dr;
        if (sub == 0 && run) {
            y[i] = out;
            if (zl) {
                const V zr = 0, zi = 0;
                cvec<V> zo;
                zo.x = out.x - (zr * dr - zi * di);
                zo.y = out.y - (zr * di + zi * dr);
                zl[i] = zo;
            }
        }
    }
}

// Backward sweep over one colour: z_i = y_i - (sum_{j in later colours} a_ij z_j) / d_i
// (ssor.h:71-78).
// LANES as in k_sweep_fwd.
template <int NS, typename V, int SUB, bool LANES>
__global__ void k_sweep_bwd(DevOp<V> op, const double* __restrict__ omega, int r0, int r1,
                            const cvec<V>* __restrict__ y, cvec<V>* __restrict__ z,
                            const SysState* __restrict__ st) {
    static_assert(!LANES || SUB == 1, "shared couplings at one lane per row");
    const int tid = threadIdx.x, sys = tid % NS, sub = (tid / NS) % SUB;
    constexpr int rpb = sweep_rpb<NS, SUB>();
    const V* ltab = LANES ? stage_value_table<NS, V>(op, dyn_lds()) : nullptr;
    const bool run = (status_mask<NS>(st, kRunning) >> sys) & 1u;
    if (!LANES && !run) return;
    const V w = V(omega[sys]);
    const int nrows = r1 - r0;
    const int nrb = (nrows + rpb - 1) / rpb;
    const RowRange R = row_range_chunk(nrb, op.shape[width_index(NS)].sweep_chunk);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = r0 + rb * rpb + tid / (NS * SUB);
        if (row >= r1) continue;
        const size_t i = size_t(row) * NS + sys;
        // Every load the row needs besides the gather goes out before it.
        const int4 sc = slice_desc(op.slc, row);
        const V ds = op.diag_s[row], me = op.me[row];
        const cvec<V> yi = y[i];
        V ar = 0, ai = 0;
        if constexpr (LANES)
            lane_block<NS, true, V>(op, ltab, sc.z, sc.w, row, z, sys, run, ar, ai);
        else
            row_gather_split<SUB, true>(op, sc.z, sc.w, row & 63, z, NS, sys, sub, ar, ai);
        V dr, di;
        cinv(ds, w * me, dr, di);
        cvec<V> out;
        out.x = yi.x - (ar * dr - ai * di);
        out.y = yi.y - (ar * di + ai * dr);
        if (sub == 0 && run) z[i] = out;
    }
}

// v = A p_hat and (r0, v); alpha = rho / (r0, v) (bicgstab.cpp:76-87).
// With ys (the forward-sweep output of p_hat) v = L p_hat + D ys.
// (A z)_row of the BiCGStab products: Eisenstat's form with ys, else the
// plain product; LANES: couplings shared among the row's lanes.
template <int NS, typename V, bool LANES>
__device__ inline cvec<V> product_row(const DevOp<V>& op, const V* __restrict__ ltab, int row, int sys, V w,
                                      const cvec<V>* __restrict__ z, const cvec<V>* __restrict__ ys, bool run) {
    if constexpr (LANES)
        return ys ? apply_row_eis_wave<NS, V>(op, ltab, row, sys, w, z, ys, run)
                  : apply_row_wave<NS, V>(op, ltab, row, sys, w, z, run);
    else
        return ys ? apply_row_eis(op, row, NS, sys, w, z, ys) : apply_row(op, row, NS, sys, w, z);
}

template <int NS, typename V, bool LANES>
__global__ void k_spmv_v(DevOp<V> op, const double* __restrict__ omega,
                         const cvec<V>* __restrict__ z, const cvec<V>* __restrict__ ys,
                         const cvec<V>* __restrict__ r0, cvec<V>* __restrict__ v,
                         const SysState* __restrict__ st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const V* ltab = LANES ? stage_value_table<NS, V>(op, dyn_lds()) : nullptr;
    const bool run = (status_mask<NS>(st, kRunning) >> sys) & 1u;
    const V w = V(omega[sys]);
    double acc[2] = {0.0, 0.0};
    const int nrb = (op.n + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = rb * rpb + tid / NS;
        if (row >= op.n || !(run || LANES)) continue;
        const size_t i = size_t(row) * NS + sys;
        const cvec<V> y = product_row<NS, V, LANES>(op, ltab, row, sys, w, z, ys, run);
        if (!run) continue;
        v[i] = y;
        const double2 d = cconjmul(to_d<V>(r0[i]), to_d<V>(y));
        acc[0] += d.x;
        acc[1] += d.y;
    }
    block_partials<NS, 2>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_spmv_v_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[2];
    if (!finish_partials<NS, 2>(part, nb, out)) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.status != kRunning) return;
    const double2 r0v = mk(out[0], out[1]);
    if (cabs(r0v) < 1e-30) {
        S.status = kBreakdown;
        S.info = 2;
        S.iter = S.iter + 1;
        S.residual = S.rnorm / S.bnorm;
        return;
    }
    S.alpha = cdiv(S.rho, r0v);
}



// t = A s_hat, (t,t), (t,s) and ||s||^2 (s was formed by the forward sweeps
// of s_hat); the finish tests ||s|| (bicgstab.cpp:94-103), then
// omega = (t,s)/(t,t) (:112-125).  With ys (the forward-sweep output of
// s_hat) t = L s_hat + D ys.
template <int NS, typename V, bool LANES>
__global__ void k_spmv_t(DevOp<V> op, const double* __restrict__ omega,
                         const cvec<V>* __restrict__ z, const cvec<V>* __restrict__ ys,
                         const cvec<V>* __restrict__ s, cvec<V>* __restrict__ t,
                         const SysState* __restrict__ st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const V* ltab = LANES ? stage_value_table<NS, V>(op, dyn_lds()) : nullptr;
    const bool run = (status_mask<NS>(st, kRunning) >> sys) & 1u;
    const V w = V(omega[sys]);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int nrb = (op.n + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = rb * rpb + tid / NS;
        if (row >= op.n || !(run || LANES)) continue;
        const size_t i = size_t(row) * NS + sys;
        const cvec<V> y = product_row<NS, V, LANES>(op, ltab, row, sys, w, z, ys, run);
        if (!run) continue;
        t[i] = y;
        const double2 td = to_d<V>(y);
        acc[0] += cabs2(td);
        const double2 sd = to_d<V>(s[i]);
        const double2 ts = cconjmul(td, sd);
        acc[1] += ts.x;
        acc[2] += ts.y;
        acc[3] += cabs2(sd);
    }
    // (t,t) and (t,s) in the three-quantity layout of the other reductions
    // (the finishing tree groups blocks by the number of quantities, so the
    // layout fixes their rounding); ||s||^2 in a region of its own.
    double tq[3] = {acc[0], acc[1], acc[2]};
    double sq[1] = {acc[3]};
    block_partials<NS, 3>(tq, part);
    block_partials<NS, 1>(sq, part + kPartSNorm);
}

template <int NS>
// ||s|| first (bicgstab.cpp:94-103): a system whose s has converged stops
// with x += alpha p_hat pending, whatever t is; then the (t, t) breakdown
// guard and omega = (t, s) / (t, t).
__global__ __launch_bounds__(kFinBlock) void k_spmv_t_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[3], sn[1];
    const bool mine = finish_partials<NS, 3>(part, nb, out);
    __syncthreads();
    finish_partials<NS, 1>(part + kPartSNorm, nb, sn);
    if (!mine) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.status != kRunning) return;
    S.snorm = sqrt(sn[0]);
    if (S.snorm / S.bnorm < S.tol) {
        S.status = kPendConvEarly;
        S.residual = S.snorm / S.bnorm;
        S.iter = S.iter + 1;
        S.info = 1;
        return;
    }
    const double tt = out[0];
    if (fabs(tt) < 1e-30) {
        S.status = kPendBreakTT;
        S.residual = S.snorm / S.bnorm;
        S.iter = S.iter + 1;
        S.info = 2;
        return;
    }
    S.omega = cdiv(mk(out[1], out[2]), mk(tt, 0.0));
}

// x += alpha p_hat + omega s_hat, r = s - omega t, ||r||, (r0, r); then the
// end-of-iteration checks and the next beta (bicgstab.cpp:127-160).
template <int NS, typename V>
__global__ void k_bicg_xr(int n, cvec<V>* __restrict__ x, const cvec<V>* __restrict__ ph,
                          const cvec<V>* __restrict__ sh, const cvec<V>* __restrict__ s,
                          const cvec<V>* __restrict__ t, const cvec<V>* __restrict__ r0,
                          cvec<V>* __restrict__ r, const SysState* __restrict__ st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    unsigned run_m = 0, pend_m = 0;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int sq = st[q].status;
        run_m |= (sq == kRunning ? 1u : 0u) << q;
        pend_m |= (sq == kPendConvEarly || sq == kPendBreakTT ? 1u : 0u) << q;
    }
    const bool run = (run_m >> sys) & 1u;
    const bool pend = (pend_m >> sys) & 1u;
    const V ar = V(st[sys].alpha.x), ai = V(st[sys].alpha.y);
    const V wr = V(st[sys].omega.x), wi = V(st[sys].omega.y);
    double acc[3] = {0.0, 0.0, 0.0};
    const int nrb = (n + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = rb * rpb + tid / NS;
        if (row >= n || !(run || pend)) continue;
        const size_t i = size_t(row) * NS + sys;
        const cvec<V> pi = ph[i];
        cvec<V> xi = x[i];
        if (pend) {
            xi.x += ar * pi.x - ai * pi.y;
            xi.y += ar * pi.y + ai * pi.x;
            x[i] = xi;
            continue;
        }
        const cvec<V> si = sh[i];
        xi.x += (ar * pi.x - ai * pi.y) + (wr * si.x - wi * si.y);
        xi.y += (ar * pi.y + ai * pi.x) + (wr * si.y + wi * si.x);
        x[i] = xi;
        const cvec<V> ss = s[i], ti = t[i];
        cvec<V> ri;
        ri.x = ss.x - (wr * ti.x - wi * ti.y);
        ri.y = ss.y - (wr * ti.y + wi * ti.x);
        r[i] = ri;
        const double2 rd = to_d<V>(ri);
        acc[0] += cabs2(rd);
        const double2 d = cconjmul(to_d<V>(r0[i]), rd);
        acc[1] += d.x;
        acc[2] += d.y;
    }
    block_partials<NS, 3>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_bicg_xr_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[3];
    if (!finish_partials<NS, 3>(part, nb, out)) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.status == kPendConvEarly) { S.status = kConvergedEarly; return; }
    if (S.status == kPendBreakTT) { S.status = kBreakdown; return; }
    if (S.status != kRunning) return;
    const int it = S.iter + 1;
    S.rnorm = sqrt(out[0]);
    const double rel = S.rnorm / S.bnorm;
    if (rel < S.tol) {
        S.status = kConverged; S.iter = it; S.residual = rel; S.info = 0;
        return;
    }
    if (cabs(S.omega) < 1e-30) {
        S.status = kBreakdown; S.iter = it; S.residual = rel; S.info = 2;
        return;
    }
    S.rho_old = S.rho;
    S.iter = it;
    S.residual = rel;
    if (it >= S.maxiter) {
        S.status = kMaxIter; S.info = 3;
        return;
    }
    S.rho = mk(out[1], out[2]);
    if (cabs(S.rho) < 1e-30) {
        S.status = kBreakdown; S.iter = it + 1; S.info = 2;
        return;
    }
    S.beta = cmul(cdiv(S.rho, S.rho_old), cdiv(S.alpha, S.omega));
}

// ---------------------------------------------------------------------------
// Plain operator application and residuals.
// ---------------------------------------------------------------------------

// y = A x, or y = b - A x when b != nullptr.  WAVE: couplings shared among
// a row's lanes (wide batches, apply_row_wave).
template <int NS, typename V, bool WAVE>
__global__ void k_spmv(DevOp<V> op, const double* __restrict__ omega,
                       const cvec<V>* __restrict__ x, const cvec<V>* __restrict__ b,
                       cvec<V>* __restrict__ y) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const V* ltab = WAVE ? stage_value_table<NS, V>(op, dyn_lds()) : nullptr;
    const V w = V(omega[sys]);
    const int nrb = (op.n + rpb - 1) / rpb;
    const int* __restrict__ order = op.order[width_index(NS)];
    const RowRange R = order ? xcd_part(nrb) : row_range(nrb);
    for (int pos = R.first; pos < R.end; pos += R.step) {
        const int rb = order ? order[pos] : pos;
        const int row = rb * rpb + tid / NS;
        if (row >= op.n) continue;
        const size_t i = size_t(row) * NS + sys;
        cvec<V> a;
        if constexpr (WAVE)
            a = apply_row_wave<NS, V>(op, ltab, row, sys, w, x);
        else
            a = apply_row(op, row, NS, sys, w, x);
        if (b) {
            const cvec<V> bi = b[i];
            a.x = bi.x - a.x;
            a.y = bi.y - a.y;
        }
        y[i] = a;
    }
}

// Bytes of LDS before a tile's staged rows: the value table, 16-byte aligned.
template <typename V>
__host__ __device__ constexpr size_t tile_table_bytes(int ntab) {
    return (size_t(ntab) * sizeof(V) + 15) / 16 * 16;
}
constexpr int kTileBlock = kTileWaves * 64;  // threads of a tiled-product workgroup (two per CU by LDS)
constexpr int kTileLoads = 4;     // staging loads in flight per thread
constexpr int kTileGroup = 8;     // consecutive tiles one XCD takes from a window

// Copies the x rows a tile stages into LDS: element i (row i / NS, system
// i % NS) of the staged block, kTileLoads per thread in flight (all row
// indices first, then all x loads, then the LDS stores).
template <int NS, typename V>
__device__ inline void stage_tile(const DevTiles& T, int s0, int ns, const cvec<V>* __restrict__ x,
                                  cvec<V>* xs) {
    const int total = ns * NS;
    for (int base = threadIdx.x; base < total; base += kTileLoads * kTileBlock) {
        int r[kTileLoads];
#pragma unroll
        for (int j = 0; j < kTileLoads; ++j) {
            const int i = base + j * kTileBlock;
            r[j] = i < total ? T.stage_rows[s0 + i / NS] : 0;
        }
        cvec<V> v[kTileLoads];
#pragma unroll
        for (int j = 0; j < kTileLoads; ++j) {
            const int i = base + j * kTileBlock;
            v[j].x = V(0);
            v[j].y = V(0);
            if (i < total) v[j] = x[size_t(r[j]) * NS + (i % NS)];
        }
#pragma unroll
        for (int j = 0; j < kTileLoads; ++j) {
            const int i = base + j * kTileBlock;
            if (i < total) xs[i] = v[j];
        }
    }
}

// y = A x (or b - A x) over the row tiles of this width (DevTiles), one tile
// per workgroup.  The workgroup copies the rows of x its tile stages into LDS
// (the NS systems of a row are NS consecutive complex values, so NS
// consecutive threads read them in one coalesced segment), then each
// wavefront computes slices of H = 64 / NS rows, the couplings of a row
// shared among its NS lanes as in apply_row_wave but every x read from LDS.
// A wavefront loads its first slice's descriptor, rows, diagonal pairs and
// couplings before the staging completes.  A row adds its entries in the
// order of its SELL blocks and its padding adds 0 * 0, so y is k_spmv's bit
// for bit.  Tiles go out in caller order in windows of kXcds x kTileGroup
// tiles, each XCD taking kTileGroup consecutive tiles of a window (workgroup
// i runs on XCD i % 8): neighbouring tiles, which stage overlapping rows,
// share an L2, and all XCDs work on one window of the mesh, whose rows the
// memory-side Infinity Cache keeps for the tiles that stage them again.
// Measured on a 1.45 M-edge mesh at 16 systems: 0.315 ms, against 0.331 ms
// for a contiguous eighth of the tiles per XCD (less fetched beyond L2, 598
// against 869 MB, but each XCD on its own region) and 0.319 ms for single
// tiles dealt round robin (1287 MB).  Results do not depend on where a
// workgroup runs.
template <int NS, typename V>
__global__ __launch_bounds__(kTileBlock) void k_spmv_tiled(DevOp<V> op, const double* __restrict__ omega,
                                                         const cvec<V>* __restrict__ x,
                                                         const cvec<V>* __restrict__ b,
                                                         cvec<V>* __restrict__ y) {
    static_assert(wave_gather_width<NS>(), "couplings shared among a row's lanes (widths 8-32)");
    constexpr int H = kSlice / NS;
    constexpr int U = 4;
    const DevTiles& T = op.tiles[width_index(NS)];
    constexpr int W = kXcds * kTileGroup;  // tiles per window
    const int wg = int(blockIdx.x);
    const int t = wg / W * W + wg % kXcds * kTileGroup + wg / kXcds % kTileGroup;
    if (t >= T.ntiles) return;  // the whole workgroup
    unsigned char* lds = dyn_lds();
    V* ltab = reinterpret_cast<V*>(lds);
    cvec<V>* xs = reinterpret_cast<cvec<V>*>(lds + tile_table_bytes<V>(op.ntab));
    const int tid = threadIdx.x, lane = tid & 63, sys = lane % NS, l = lane / NS, gl = lane - sys;
    const V w = V(omega[sys]);
    const int lb = T.lbits;
    const unsigned lm = (1u << lb) - 1u;
    const int c1 = T.slice_off[t + 1];
    int c = T.slice_off[t] + (tid >> 6);
    // The first slice's loads, which do not depend on x: descriptor, rows,
    // diagonal pairs and the first two chunks of couplings.
    int4 d = make_int4(0, 0, 0, 0);
    int2 so = make_int2(-1, 0);
    V dr = 0, dm = 0;
    unsigned mine0 = 0u, mine1 = 0u;
    auto slice_loads = [&](int cc) {
        d = T.slc[__builtin_amdgcn_readfirstlane(cc)];
        so = T.slot[cc * H + l];
        if (so.x >= 0) {
            dr = op.diag_s[so.x];
            dm = op.me[so.x];
        }
        mine0 = sys < d.y ? T.words[d.x + sys * H + l] : 0u;
        mine1 = NS + sys < d.y ? T.words[d.x + (NS + sys) * H + l] : 0u;
    };
    if (c < c1) slice_loads(c);
    for (int i = tid; i < op.ntab; i += kTileBlock) ltab[i] = op.vtab[i];
    const int s0 = T.stage_off[t];
    stage_tile<NS, V>(T, s0, T.stage_off[t + 1] - s0, x, xs);
    __syncthreads();
    for (; c < c1; c += kTileWaves) {
        const int off = __builtin_amdgcn_readfirstlane(d.x), wd = __builtin_amdgcn_readfirstlane(d.y);
        V ar = 0, ai = 0;
        for (int k0 = 0; k0 < wd; k0 += NS) {
            const unsigned mine = k0 == 0        ? mine0
                                  : k0 == NS     ? mine1
                                  : k0 + sys < wd ? T.words[off + (k0 + sys) * H + l]
                                                  : 0u;
            const int n = min(NS, wd - k0);
            for (int j = 0; j < n; j += U) {
                unsigned q[U];
#pragma unroll
                for (int u = 0; u < U; ++u) q[u] = lane_get(mine, gl + j + u);
                cvec<V> xv[U];
                V a[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    xv[u].x = V(0);
                    xv[u].y = V(0);
                    a[u] = V(0);
                    if (j + u < n) {
                        xv[u] = xs[(q[u] & lm) * NS + sys];
                        a[u] = ltab[q[u] >> lb];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ar += a[u] * xv[u].x;
                    ai += a[u] * xv[u].y;
                }
            }
        }
        if (so.x >= 0) {
            const V di = w * dm;
            const cvec<V> xi = xs[so.y * NS + sys];
            cvec<V> out;
            out.x = ar + (dr * xi.x - di * xi.y);
            out.y = ai + (dr * xi.y + di * xi.x);
            const size_t i = size_t(so.x) * NS + sys;
            if (b) {
                const cvec<V> bi = b[i];
                out.x = bi.x - out.x;
                out.y = bi.y - out.y;
            }
            y[i] = out;
        }
        // The next slice of this wavefront (none when the tile holds at most
        // one slice per wavefront, as the tiles are built).
        if (c + kTileWaves < c1) slice_loads(c + kTileWaves);
    }
}

// ||A x - b|| / ||b|| in FP64 (forward_solver_fv.cpp:843-848, 786-790); with
// r != nullptr also stores r = b - A x in FP64 (iterative refinement).
template <int NS>
__global__ void k_true_res(DevOp<double> op, const double* __restrict__ omega,
                           const double2* __restrict__ x, const double2* __restrict__ b,
                           double2* __restrict__ r, SysState* st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const double w = omega[sys];
    double acc[2] = {0.0, 0.0};
    const int nrb = (op.n + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = rb * rpb + tid / NS;
        if (row >= op.n) continue;
        const size_t i = size_t(row) * NS + sys;
        const double2 ax = apply_row(op, row, NS, sys, w, x);
        const double2 bi = b[i];
        const double2 d = mk(bi.x - ax.x, bi.y - ax.y);
        if (r) r[i] = d;
        acc[0] += cabs2(d);
        acc[1] += cabs2(bi);
    }
    block_partials<NS, 2>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_true_res_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[2];
    if (!finish_partials<NS, 2>(part, nb, out)) return;
    const int tid = threadIdx.x;
    double bn = sqrt(out[1]);
    if (bn < 1e-30) bn = 1.0;
    st[tid].true_bnorm = bn;
    st[tid].true_res = sqrt(out[0]) / bn;
}

// Least-squares scale of an initial guess: y = A x0, gamma = (y, b) / (y, y)
// per system, so that ||b - gamma A x0|| is minimal.
template <int NS>
__global__ void k_guess_dots(DevOp<double> op, const double* __restrict__ omega,
                             const double2* __restrict__ x, const double2* __restrict__ b, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const double w = omega[sys];
    double acc[3] = {0.0, 0.0, 0.0};
    const int nrb = (op.n + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int row = rb * rpb + tid / NS;
        if (row >= op.n) continue;
        const size_t i = size_t(row) * NS + sys;
        const double2 y = apply_row(op, row, NS, sys, w, x);
        const double2 yb = cconjmul(y, b[i]);
        acc[0] += yb.x;
        acc[1] += yb.y;
        acc[2] += cabs2(y);
    }
    block_partials<NS, 3>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_guess_dots_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[3];
    if (!finish_partials<NS, 3>(part, nb, out)) return;
    const int tid = threadIdx.x;
    st[tid].guess_scale = out[2] > 0.0 ? cdiv(mk(out[0], out[1]), mk(out[2], 0.0)) : mk(0.0, 0.0);
}

// x *= guess_scale of its system.
template <int NS>
__global__ void k_scale_guess(int npad, const SysState* __restrict__ st, double2* __restrict__ x) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        x[t] = cmul(st[t % NS].guess_scale, x[t]);
    }
}

// Misfit gradient in ln sigma, summed over systems (compute_sensitivity,
// forward_solver_fv.cpp:989-1059): g_c = sum_s Re(i w_s sigma_c a_s) with
// a_s = sum_{e in c} lambda_s[e] E_s[e] w_e (lambda unconjugated); one thread
// per cell, fields system-major in caller order.
__global__ void k_sensitivity(int nc, int n, int nsys, const int* __restrict__ cptr,
                              const int* __restrict__ cedge, const double* __restrict__ wgt,
                              const double* __restrict__ sigma, const double* __restrict__ omega,
                              const double2* __restrict__ E, const double2* __restrict__ lam,
                              double* __restrict__ g) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int s = 0; s < nsys; ++s) {
            const double2* Es = E + size_t(s) * n;
            const double2* Ls = lam + size_t(s) * n;
            double2 a = mk(0.0, 0.0);
            for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                const int e = cedge[k];
                const double w = wgt[e];
                if (w == 0.0) continue;
                const double2 p = cmul(Ls[e], Es[e]);
                a.x += p.x * w;
                a.y += p.y * w;
            }
            // Re(i w sigma a) = -w sigma Im(a)
            acc += -omega[s] * sigma[c] * a.y;
        }
        g[c] = acc;
    }
}

void launch_sensitivity(hipStream_t s, int nc, int n, int nsys, const int* cptr, const int* cedge,
                        const double* wgt, const double* sigma, const double* omega, const double2* E,
                        const double2* lam, double* g) {
    const int blocks = std::max(1, std::min((nc + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_sensitivity, dim3(blocks), dim3(kBlock), 0, s, nc, n, nsys, cptr, cedge, wgt,
                       sigma, omega, E, lam, g);
}

// ---------------------------------------------------------------------------
// Vector utilities.
// ---------------------------------------------------------------------------

// Move systems between batch widths: narrow[row*nb + j] = wide[row*na + map[j]]
// (to_narrow) or the reverse, for map[j] >= 0.
template <typename V>
__global__ void k_regroup(int rows, int na, int nb, const int* __restrict__ map,
                          cvec<V>* __restrict__ wide, cvec<V>* __restrict__ narrow, int to_narrow) {
    const size_t total = size_t(rows) * nb;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int row = int(t / nb), j = int(t % nb);
        const int src = map[j];
        if (to_narrow) {
            cvec<V> z;
            z.x = 0;
            z.y = 0;
            if (src >= 0) z = wide[size_t(row) * na + src];
            narrow[t] = z;
        } else if (src >= 0) {
            wide[size_t(row) * na + src] = narrow[t];
        }
    }
}
template <int NS, typename V>
__global__ void k_fill_zero(size_t count, cvec<V>* __restrict__ x) {
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < count;
         t += size_t(gridDim.x) * blockDim.x) {
        cvec<V> z;
        z.x = 0;
        z.y = 0;
        x[t] = z;
    }
}

// Zero the solution of systems whose flag is set (divergence guard,
// forward_solver_fv.cpp:749-752, 821-826).
template <int NS>
__global__ void k_zero_systems(int npad, const int* __restrict__ flags, double2* __restrict__ x) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        if (flags[t % NS]) x[t] = mk(0.0, 0.0);
    }
}

// dst = src (0 when src is null) on the systems whose flag is set.
template <int NS>
__global__ void k_copy_systems(int npad, const int* __restrict__ flags, const double2* __restrict__ src,
                               double2* __restrict__ dst) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        if (flags[t % NS]) dst[t] = src ? src[t] : mk(0.0, 0.0);
    }
}

// E[coarse] = 0.5 (E[child1] + E[child2]) for one independent level of
// hanging pairs (forward_solver_fv.cpp:258-261).
template <int NS, typename V>
__global__ void k_interp_hanging(int npairs, const int* __restrict__ tri, cvec<V>* __restrict__ x) {
    const size_t total = size_t(npairs) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int p = int(t / NS), sys = int(t % NS);
        const int c = tri[3 * p], a = tri[3 * p + 1], b = tri[3 * p + 2];
        const cvec<V> xa = x[size_t(a) * NS + sys], xb = x[size_t(b) * NS + sys];
        cvec<V> o;
        o.x = V(0.5) * (xa.x + xb.x);
        o.y = V(0.5) * (xa.y + xb.y);
        x[size_t(c) * NS + sys] = o;
    }
}

// E = E_pri + E_sec, then Dirichlet rows take rhs_total (E_bc)
// (forward_solver_fv.cpp:357-361).
template <int NS>
__global__ void k_forward_total(int npad, const unsigned char* __restrict__ bc,
                                const double2* __restrict__ epri, const double2* __restrict__ b,
                                double2* __restrict__ x) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int row = int(t / NS);
        if (bc[row]) {
            x[t] = b[t];
        } else {
            const double2 a = epri[t], c = x[t];
            x[t] = mk(a.x + c.x, a.y + c.y);
        }
    }
}

// Constrained (identity) rows of a supplied initial guess take their exact
// value x_i = b_i.
template <int NS>
__global__ void k_fix_constrained(int npad, const unsigned char* __restrict__ cons,
                                  const double2* __restrict__ b, double2* __restrict__ x) {
    const size_t total = size_t(npad) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x)
        if (cons[t / NS]) x[t] = b[t];
}

// x += d (FP64 accumulate of an FP32 correction).
template <int NS>
__global__ void k_add_correction(size_t count, const float2* __restrict__ d, double2* __restrict__ x) {
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < count;
         t += size_t(gridDim.x) * blockDim.x) {
        const float2 a = d[t];
        const double2 c = x[t];
        x[t] = mk(c.x + double(a.x), c.y + double(a.y));
    }
}

// FP32 copy of a (scaled) FP64 vector: dst = src * scale[sys].
template <int NS>
__global__ void k_demote(size_t count, const double2* __restrict__ src, const double* __restrict__ scale,
                         float2* __restrict__ dst) {
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < count;
         t += size_t(gridDim.x) * blockDim.x) {
        const double2 a = src[t];
        const double sc = scale[t % NS];
        dst[t] = make_float2(float(a.x * sc), float(a.y * sc));
    }
}

// x += d * scale[sys] (undo the right-hand-side scaling of a correction).
template <int NS>
__global__ void k_add_scaled(size_t count, const float2* __restrict__ d, const double* __restrict__ scale,
                             double2* __restrict__ x) {
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < count;
         t += size_t(gridDim.x) * blockDim.x) {
        const float2 a = d[t];
        const double sc = scale[t % NS];
        const double2 c = x[t];
        x[t] = mk(c.x + double(a.x) * sc, c.y + double(a.y) * sc);
    }
}

// ---------------------------------------------------------------------------
// Divergence correction (divcorr_poisson_, forward_solver_fv.cpp:623-722),
// FP64 throughout.  Node n's incident edges are nt_ptr[n]..nt_ptr[n+1] in
// ascending caller edge order; nt_edge holds device rows, nt_val the G value.
// ---------------------------------------------------------------------------

// phi0 = G^T b / (i omega) over non-dead edges (compute_phi0, :893-908).
template <int NS>
__global__ void k_phi0(int nn, const int* __restrict__ nt_ptr, const int* __restrict__ nt_edge,
                       const double* __restrict__ nt_val, const unsigned char* __restrict__ dead,
                       const double* __restrict__ omega, const double2* __restrict__ b,
                       double2* __restrict__ phi0) {
    const size_t total = size_t(nn) * NS;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int v = int(t / NS), sys = int(t % NS);
        double2 acc = mk(0.0, 0.0);
        for (int k = nt_ptr[v]; k < nt_ptr[v + 1]; ++k) {
            const int e = nt_edge[k];
            if (dead[e]) continue;
            const double g = nt_val[k];
            const double2 bv = b[size_t(e) * NS + sys];
            acc.x += g * bv.x;
            acc.y += g * bv.y;
        }
        phi0[t] = cdiv(acc, mk(0.0, omega[sys]));
    }
}

// Bit s set when DivCorr (pcg = false) or its PCG (pcg = true) of system s
// is running; scalar loads, as status_mask.
template <int NS>
__device__ inline unsigned dc_mask(const SysState* __restrict__ st, bool pcg) {
    unsigned m = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) m |= ((pcg ? st[s].pcg_status : st[s].dc_status) == kRunning ? 1u : 0u) << s;
    return m;
}

__global__ void k_dc_begin(SysState* st, unsigned mask) {
    const int s = threadIdx.x;
    if (s >= kMaxSys) return;
    st[s].dc_status = (mask & (1u << s)) ? kRunning : kIdle;
    st[s].pcg_status = kIdle;
}

// The DivCorr PCG runs in V: double everywhere, float (complex64 vectors,
// FP32 Laplacian values) for the interleaved corrections of the mixed
// precision path.  Reductions accumulate in double in both; with V = double
// every expression below is the FP64 one, operation for operation.
template <typename V>
__device__ inline cvec<V> mkv(V a, V b) {
    cvec<V> r;
    r.x = a;
    r.y = b;
    return r;
}

// divJ = G^T (Me_sigma E) - phi0, zero on Dirichlet nodes, ||divJ||; PCG start
// r = divJ, z = r / diag, p = z, phi = 0, rz = (r, z) (:636-677).  divJ is
// formed in double from the FP64 field; its norm (the DivCorr convergence
// measure) is that of the double value, the PCG start is divJ rounded to V.
template <int NS, typename V>
__global__ void k_dc_divj(int nn, const int* __restrict__ nt_ptr, const int* __restrict__ nt_edge,
                          const double* __restrict__ nt_val, const double* __restrict__ ms,
                          const unsigned char* __restrict__ bc_node, const V* __restrict__ ldiag,
                          const double2* __restrict__ E, const double2* __restrict__ phi0,
                          cvec<V>* __restrict__ r, cvec<V>* __restrict__ z, cvec<V>* __restrict__ p,
                          cvec<V>* __restrict__ phi, const SysState* __restrict__ st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const bool run = (dc_mask<NS>(st, false) >> sys) & 1u;
    double acc[3] = {0.0, 0.0, 0.0};
    const int nrb = (nn + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int v = rb * rpb + tid / NS;
        if (v >= nn || !run) continue;
        const size_t i = size_t(v) * NS + sys;
        double2 a = mk(0.0, 0.0);
        for (int k = nt_ptr[v]; k < nt_ptr[v + 1]; ++k) {
            const int e = nt_edge[k];
            const double2 ev = E[size_t(e) * NS + sys];
            const double m = ms[e];
            const double g = nt_val[k];
            a.x += g * (m * ev.x);
            a.y += g * (m * ev.y);
        }
        if (phi0) {
            const double2 f = phi0[i];
            a.x -= f.x;
            a.y -= f.y;
        }
        if (bc_node[v]) a = mk(0.0, 0.0);
        const cvec<V> av = to_v<V>(a);
        r[i] = av;
        const V d = ldiag[v];
        const cvec<V> zi = (fabs(d) > V(1e-30)) ? mkv<V>(av.x / d, av.y / d) : av;
        z[i] = zi;
        p[i] = zi;
        phi[i] = mkv<V>(V(0), V(0));
        acc[0] += cabs2(a);
        const double2 rz = cconjmul(to_d<V>(av), to_d<V>(zi));
        acc[1] += rz.x;
        acc[2] += rz.y;
    }
    block_partials<NS, 3>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_dc_divj_fin(int dc_iter, double dc_tol, int cg_max, double pcg_tol,
                                                            SysState* st, const double* __restrict__ part, int nb) {
    double out[3];
    if (!finish_partials<NS, 3>(part, nb, out)) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.dc_status != kRunning) return;
    const double dn = sqrt(out[0]);
    if (dc_iter == 0) S.div_norm0 = dn;
    S.div_norm = dn;
    const double rel = (S.div_norm0 > 1e-30) ? dn / S.div_norm0 : 0.0;
    if ((rel < dc_tol && dc_iter > 0) || dn < 1e-12) {
        S.dc_status = kConverged;
        S.pcg_status = kIdle;
        return;
    }
    S.rz = mk(out[1], out[2]);
    S.pcg_status = kRunning;
    S.pcg_iter = 0;
    S.pcg_max = cg_max;
    S.pcg_tol = pcg_tol;
}

// (L p)_v with the row's NS lanes sharing its entries (wide batches): lane j
// loads entry kb + j of each chunk of NS entries and passes its column and
// value to the row's lanes through ds_bpermute, as the SELL kernels do.  The
// entries are added in increasing k, as in k_dc_lp's per-lane loop, so the
// sums are the same bit for bit.  A lane whose system is not running (run =
// false) passes its entries on and gathers nothing.
template <int NS, typename V>
__device__ inline cvec<V> dc_row_lanes(int v, const int* __restrict__ lptr, const int* __restrict__ lcol,
                                       const V* __restrict__ lval, const cvec<V>* __restrict__ p, int sys,
                                       bool run) {
    constexpr int U = 8;
    static_assert(NS % U == 0, "a batch stays within the row's lanes");
    const int gl = (threadIdx.x & 63) - sys;
    const int k0 = lptr[v], k1 = lptr[v + 1];
    cvec<V> a = mkv<V>(V(0), V(0));
    for (int kb = k0; kb < k1; kb += NS) {
        const int k = kb + sys;
        int c = 0;
        V l = V(0);
        if (k < k1) {
            c = lcol[k];
            l = lval[k];
        }
        const int n = min(NS, k1 - kb);
        for (int j = 0; j < n; j += U) {
            int cc[U];
            V ll[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cc[u] = lane_get(c, gl + j + u);
                ll[u] = lane_get(l, gl + j + u);
            }
            cvec<V> pv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                pv[u] = mkv<V>(V(0), V(0));
                if (run && j + u < n) pv[u] = p[size_t(cc[u]) * NS + sys];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j + u < n) {
                    a.x += ll[u] * pv[u].x;
                    a.y += ll[u] * pv[u].y;
                }
            }
        }
    }
    return a;
}

// Ap = L p, (p, Ap); alpha = rz / pAp (:681-692).  LANES: the entries of a
// row shared among its lanes (dc_row_lanes).
template <int NS, bool LANES, typename V>
__global__ void k_dc_lp(int nn, const int* __restrict__ lptr, const int* __restrict__ lcol,
                        const V* __restrict__ lval, const cvec<V>* __restrict__ p,
                        cvec<V>* __restrict__ ap, const SysState* __restrict__ st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const bool run = (dc_mask<NS>(st, true) >> sys) & 1u;
    double acc[2] = {0.0, 0.0};
    const int nrb = (nn + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int v = rb * rpb + tid / NS;
        if (v >= nn || !(run || LANES)) continue;
        const size_t i = size_t(v) * NS + sys;
        if constexpr (LANES) {
            const cvec<V> a = dc_row_lanes<NS, V>(v, lptr, lcol, lval, p, sys, run);
            if (!run) continue;
            ap[i] = a;
            const double2 d = cconjmul(to_d<V>(p[i]), to_d<V>(a));
            acc[0] += d.x;
            acc[1] += d.y;
            continue;
        }
        // Entries in increasing k, indices and values of 8 entries loaded
        // before their p entries.
        cvec<V> a = mkv<V>(V(0), V(0));
        const int k1 = lptr[v + 1];
        for (int k = lptr[v]; k < k1; k += 8) {
            int c[8];
            V l[8];
            cvec<V> pv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                c[u] = 0;
                l[u] = V(0);
                if (k + u < k1) {
                    c[u] = lcol[k + u];
                    l[u] = lval[k + u];
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < k1) pv[u] = p[size_t(c[u]) * NS + sys];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (k + u < k1) {
                    a.x += l[u] * pv[u].x;
                    a.y += l[u] * pv[u].y;
                }
            }
        }
        ap[i] = a;
        const double2 d = cconjmul(to_d<V>(p[i]), to_d<V>(a));
        acc[0] += d.x;
        acc[1] += d.y;
    }
    block_partials<NS, 2>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_dc_lp_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[2];
    if (!finish_partials<NS, 2>(part, nb, out)) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.pcg_status != kRunning) return;
    const double2 pap = mk(out[0], out[1]);
    if (cabs(pap) < 1e-30) {
        S.pcg_status = kConverged;
        return;
    }
    S.pcg_alpha = cdiv(S.rz, pap);
}

// phi += alpha p, r -= alpha Ap, ||r||; z = r / diag, (r, z);
// stop when ||r|| < 1e-10 ||divJ||, else beta = rz_new / (rz + 1e-30) (:693-711).
template <int NS, typename V>
__global__ void k_dc_update(int nn, const V* __restrict__ ldiag, const cvec<V>* __restrict__ p,
                            const cvec<V>* __restrict__ ap, cvec<V>* __restrict__ phi,
                            cvec<V>* __restrict__ r, const SysState* __restrict__ st, double* part) {
    const int tid = threadIdx.x, sys = tid % NS;
    const int rpb = rows_per_block<NS>();
    const bool run = (dc_mask<NS>(st, true) >> sys) & 1u;
    const V ax = V(st[sys].pcg_alpha.x), ay = V(st[sys].pcg_alpha.y);
    double acc[3] = {0.0, 0.0, 0.0};
    const int nrb = (nn + rpb - 1) / rpb;
    const RowRange R = row_range(nrb);
    for (int rb = R.first; rb < R.end; rb += R.step) {
        const int v = rb * rpb + tid / NS;
        if (v >= nn || !run) continue;
        const size_t i = size_t(v) * NS + sys;
        const cvec<V> pi = p[i], api = ap[i];
        const cvec<V> ap_ = mkv<V>(ax * pi.x - ay * pi.y, ax * pi.y + ay * pi.x);
        const cvec<V> aap = mkv<V>(ax * api.x - ay * api.y, ax * api.y + ay * api.x);
        cvec<V> f = phi[i];
        f.x += ap_.x;
        f.y += ap_.y;
        phi[i] = f;
        cvec<V> ri = r[i];
        ri.x -= aap.x;
        ri.y -= aap.y;
        r[i] = ri;
        // z = D^-1 r enters the (r, z) product only: k_dc_p forms it again
        // from r with the same division instead of a stored copy.
        const V d = ldiag[v];
        const cvec<V> zi = (fabs(d) > V(1e-30)) ? mkv<V>(ri.x / d, ri.y / d) : ri;
        const double2 rd = to_d<V>(ri);
        acc[0] += cabs2(rd);
        const double2 rz = cconjmul(rd, to_d<V>(zi));
        acc[1] += rz.x;
        acc[2] += rz.y;
    }
    block_partials<NS, 3>(acc, part);
}

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_dc_update_fin(SysState* st, const double* __restrict__ part, int nb) {
    double out[3];
    if (!finish_partials<NS, 3>(part, nb, out)) return;
    const int tid = threadIdx.x;
    SysState& S = st[tid];
    if (S.pcg_status != kRunning) return;
    S.pcg_iter += 1;
    S.pcg_rnorm = sqrt(out[0]);
    if (S.pcg_rnorm < S.pcg_tol * S.div_norm || S.pcg_iter >= S.pcg_max) {
        S.pcg_status = kConverged;
        return;
    }
    const double2 rzn = mk(out[1], out[2]);
    S.pcg_beta = cdiv(rzn, mk(S.rz.x + 1e-30, S.rz.y));
    S.rz = rzn;
}

// p = z + beta p with z = D^-1 r formed here (the Jacobi preconditioner of
// k_dc_update, same division).
template <int NS, typename V>
__global__ void k_dc_p(int nn, const V* __restrict__ ldiag, const cvec<V>* __restrict__ r,
                       cvec<V>* __restrict__ p, const SysState* __restrict__ st) {
    const size_t total = size_t(nn) * NS;
    const unsigned run = dc_mask<NS>(st, true);
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int sys = int(t % NS);
        if (!((run >> sys) & 1u)) continue;
        const V bx = V(st[sys].pcg_beta.x), by = V(st[sys].pcg_beta.y);
        const cvec<V> pt = p[t];
        const cvec<V> bp = mkv<V>(bx * pt.x - by * pt.y, bx * pt.y + by * pt.x);
        const cvec<V> ri = r[t];
        const V d = ldiag[t / NS];
        const cvec<V> zi = (fabs(d) > V(1e-30)) ? mkv<V>(ri.x / d, ri.y / d) : ri;
        p[t] = mkv<V>(zi.x + bp.x, zi.y + bp.y);
    }
}

// E -= G phi on edges outside the DivCorr Dirichlet set (:714-720).
template <int NS, typename V>
__global__ void k_dc_correct(int npad, const int* __restrict__ g_col, const double* __restrict__ g_val,
                             const unsigned char* __restrict__ bc_edge,
                             const cvec<V>* __restrict__ phi, double2* __restrict__ E,
                             const SysState* __restrict__ st) {
    const size_t total = size_t(npad) * NS;
    const unsigned run = dc_mask<NS>(st, false);
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int row = int(t / NS), sys = int(t % NS);
        if (!((run >> sys) & 1u) || bc_edge[row]) continue;
        const int a = g_col[2 * row], b = g_col[2 * row + 1];
        if (a < 0) continue;  // padding row
        const double2 pa = to_d<V>(phi[size_t(a) * NS + sys]), pb = to_d<V>(phi[size_t(b) * NS + sys]);
        const double ga = g_val[2 * row], gb = g_val[2 * row + 1];
        double2 gr;
        gr.x = ga * pa.x + gb * pb.x;
        gr.y = ga * pa.y + gb * pb.y;
        double2 e = E[t];
        E[t] = mk(e.x - gr.x, e.y - gr.y);
    }
}

// ---------------------------------------------------------------------------
// Sigma update on the device: values only, the sparsity patterns stay.  Every
// sum runs in the host builders' order and products are not contracted into
// FMAs, so the results equal the host path's bit for bit.
// ---------------------------------------------------------------------------

// Per device row r (edge e = perm[r]): the operator's Me (edge_mass,
// operators.cpp:481-518; 0 on massless, adjacency-free, constrained and
// padding rows) in FP64 and FP32, and the DivCorr weight Me_sigma
// (edge_mass_divcorr; 0 on dead edges).  *err |= 1 when a weight crosses the
// 1e-30 threshold that decided the Laplacian pattern (in_pattern[r]).
__global__ void k_sigma_edges(int npad, int nc, const int* __restrict__ perm, const int* __restrict__ adj_ptr,
                              const int* __restrict__ adj_cell, const double* __restrict__ area,
                              const double* __restrict__ len, const unsigned char* __restrict__ massless,
                              const unsigned char* __restrict__ dead, const unsigned char* __restrict__ constrained,
                              const unsigned char* __restrict__ dc_dead, const unsigned char* __restrict__ in_pattern,
                              const double* __restrict__ sigma, double* __restrict__ me, float* __restrict__ me32,
                              double* __restrict__ ms, int* __restrict__ err) {
#pragma clang fp contract(off)
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < npad; r += gridDim.x * blockDim.x) {
        const int e = perm[r];
        double m = 0.0, w = 0.0;
        if (e >= 0) {
            const int a = adj_ptr[e], b = adj_ptr[e + 1];
            double sum = 0.0;
            int count = 0;
            for (int k = a; k < b; ++k) {
                const int c = adj_cell[k];
                if (c >= 0 && c < nc) {
                    sum += sigma[c];
                    ++count;
                }
            }
            const double avg = count > 0 ? sum / count : 1e-8;
            if (b > a && !massless[e] && !constrained[r]) {
                const double hodge = area[e] * len[e];
                m = avg * hodge;
            }
            if (!dead[e] && !dc_dead[r]) w = avg * area[e] * len[e];
            if ((w >= 1e-30) != (in_pattern[r] != 0)) atomicOr(err, 1);
        }
        me[r] = m;
        me32[r] = float(m);
        ms[r] = w;
    }
}

// Per node v: the values of row v of L = G^T diag(Me_sigma) G
// (div_laplacian, operators.cpp:952-1019: the diagonal adds g w g over the
// incident edges in ascending edge order, an off-diagonal entry is the one
// edge joining the two nodes) with the Dirichlet rewrite and the Jacobi
// diagonal taken before it (compute_bc_flags_, forward_solver_fv.cpp:581-606).
// ledge[k] is the device row of the edge of off-diagonal entry k.
__global__ void k_sigma_laplacian(int nn, const int* __restrict__ nt_ptr, const int* __restrict__ nt_edge,
                                  const double* __restrict__ nt_val, const double* __restrict__ ms,
                                  const int* __restrict__ lptr, const int* __restrict__ lcol,
                                  const int* __restrict__ ledge, const int* __restrict__ g_col,
                                  const double* __restrict__ g_val, const unsigned char* __restrict__ bc_node,
                                  double* __restrict__ lval, double* __restrict__ ldiag) {
#pragma clang fp contract(off)
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nn; v += gridDim.x * blockDim.x) {
        double d = 0.0;
        for (int q = nt_ptr[v]; q < nt_ptr[v + 1]; ++q) {
            const double w = ms[nt_edge[q]];
            if (w < 1e-30) continue;
            const double g = nt_val[q];
            d += g * w * g;
        }
        const bool bv = bc_node[v] != 0;
        bool has_diag = false;
        for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
            const int u = lcol[k];
            double val;
            if (u == v) {
                val = bv ? 1.0 : d;
                has_diag = true;
            } else if (bv || bc_node[u]) {
                val = 0.0;
            } else {
                const int r = ledge[k];
                const double w = ms[r];
                const double gv = g_col[2 * r] == v ? g_val[2 * r] : g_val[2 * r + 1];
                const double gu = g_col[2 * r] == u ? g_val[2 * r] : g_val[2 * r + 1];
                val = gv * w * gu;
            }
            lval[k] = val;
        }
        ldiag[v] = (bv || !has_diag) ? 1.0 : d;
    }
}

// Perturbation right-hand side of the Gauss-Newton product
// (build_perturbation_rhs, forward_solver_fv.cpp:945-984) for edge e of
// system s (caller order, t = s * n + e):
//   rhs = -sum over cells c containing e with |dsig_c| >= 1e-30 of
//         (i w)(dsig_c dMe_ce) E_bg[e],  dMe_ce = L^3 savg^2 / (n_adj sigma_c^2),
// savg the harmonic mean of sigma over the edge's adjacent cells.  Cells in
// ascending order (e_cells), complex products written out in the host's
// operation order, nothing contracted: equal to the host's values.
__global__ void k_perturbation_rhs(int n, int nsys, int nc, const int* __restrict__ e_ptr,
                                   const int* __restrict__ e_cells, const int* __restrict__ adj_ptr,
                                   const int* __restrict__ adj_cell, const double* __restrict__ len,
                                   const double* __restrict__ sigma, const double* __restrict__ dsig,
                                   const double* __restrict__ omega, const double2* __restrict__ ebg,
                                   double2* __restrict__ rhs) {
#pragma clang fp contract(off)
    const size_t total = size_t(n) * nsys;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int s = int(t / n), e = int(t % n);
        double ax = 0.0, ay = 0.0;
        const int a = e_ptr[e], b = e_ptr[e + 1];
        if (a < b) {
            double sum_inv = 0.0;
            int n_adj = 0;
            for (int k = adj_ptr[e]; k < adj_ptr[e + 1]; ++k) {
                const int c = adj_cell[k];
                if (c >= 0 && c < nc && sigma[c] > 1e-30) {
                    sum_inv += 1.0 / sigma[c];
                    ++n_adj;
                }
            }
            const double savg = (n_adj > 0 && sum_inv > 1e-30) ? n_adj / sum_inv : 1e-8;
            const double L = len[e];
            const double vol = L * L * L;
            const double w = omega[s];
            const double2 E = ebg[t];
            for (int k = a; k < b; ++k) {
                const int c = e_cells[k];
                const double ds = dsig[c];
                if (fabs(ds) < 1e-30) continue;
                const double sc = sigma[c];
                const double dme = vol * savg * savg / (n_adj * sc * sc);
                const double av = ds * dme;
                // (i w) * (av + 0i), then * E
                const double t1r = 0.0 * av - w * 0.0;
                const double t1i = 0.0 * 0.0 + w * av;
                const double pr = t1r * E.x - t1i * E.y;
                const double pi = t1r * E.y + t1i * E.x;
                ax -= pr;
                ay -= pi;
            }
        }
        rhs[t] = mk(ax, ay);
    }
}

// ---------------------------------------------------------------------------
// Launchers (explicit instantiation per batch width).
// ---------------------------------------------------------------------------
namespace {
inline int grid_for(size_t work_items, int cap) {
    const size_t b = (work_items + kBlock - 1) / kBlock;
    return int(b < size_t(cap) ? (b > 0 ? b : 1) : cap);
}
// Grid of a row kernel: one block per row block up to cap, rounded up to a
// multiple of the XCD count (so every XCD gets the same number of blocks;
// surplus blocks find no rows).
inline int grid_rows(int rows, int rpb, int cap) {
    int b = (rows + rpb - 1) / rpb;
    b = b < cap ? b : cap;
    b = (b + kXcds - 1) / kXcds * kXcds;
    return b > 0 ? b : kXcds;
}
template <int NS>
inline int row_grid(int rows, int cap = kMaxBlocks) {
    return grid_rows(rows, kBlock / NS, cap);
}
constexpr int kStreamCap = 4096;
static_assert(kMaxBlocks % kXcds == 0 && kStreamCap % kXcds == 0, "grid caps must split over the XCDs");
}  // namespace

template <int NS, typename V>
void Launch<NS, V>::pack(hipStream_t s, int npad, int n, int nsys, const int* perm,
                         const double2* src, cvec<V>* dst) {
    hipLaunchKernelGGL((k_pack<NS, V>), dim3(grid_for(size_t(npad) * NS, kStreamCap)), dim3(kBlock),
                       0, s, npad, n, nsys, perm, src, dst);
}
template <int NS, typename V>
void Launch<NS, V>::unpack(hipStream_t s, int npad, int n, int nsys, const int* perm,
                           const cvec<V>* src, double2* dst) {
    hipLaunchKernelGGL((k_unpack<NS, V>), dim3(grid_for(size_t(npad) * NS, kStreamCap)),
                       dim3(kBlock), 0, s, npad, n, nsys, perm, src, dst);
}
template <int NS, typename V>
void Launch<NS, V>::bicg_init(hipStream_t s, const DevOp<V>& op, const double* omega,
                              const cvec<V>* b, const cvec<V>* x, cvec<V>* r, cvec<V>* r0,
                              SysState* st, double* part, int x_zero) {
    const int g = row_grid<NS>(op.n);
    hipLaunchKernelGGL((k_bicg_init<NS, V>), dim3(g), dim3(kBlock), 0, s, op, omega, b, x, r, r0, st,
                       part, x_zero);
    hipLaunchKernelGGL((k_bicg_init_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
// Sweep launch at lanes-per-row SUB; widths where NS * SUB > 32 fall back to
// one lane per row.
template <int NS, int SUB>
constexpr bool sub_ok() { return NS * SUB <= 32; }

// Workgroups of a sweep launch over `rows` rows under the operator's shape.
template <int NS, int SUB, typename V>
int sweep_grid(const DevOp<V>& op, int rows) {
    const LaunchShape& sh = op.shape[width_index(NS)];
    constexpr int rpb = sweep_rpb<NS, SUB>();
    if (sh.sweep_chunk <= 1) return grid_rows(rows, rpb, sh.sweep_cap);
    const int nrb = (rows + rpb - 1) / rpb;
    return grid_rows(nrb, sh.sweep_chunk, 1 << 30);
}

// Couplings shared among a row's lanes at this width (launch shape wave = 1;
// the sweeps only at one lane per row), and the dynamic LDS such a launch
// needs: the value table when the width gathers from the packed blocks.
template <int NS, typename V>
bool lanes_on(const DevOp<V>& op) {
    return wave_gather_width<NS>() && op.shape[width_index(NS)].wave;
}
template <int NS, typename V>
size_t lanes_lds(const DevOp<V>& op) {
    return op.vtab && op.shape[width_index(NS)].packed ? op.ntab * sizeof(V) : 0;
}

template <int NS, typename V, int SUB, bool LANES>
void sweep_fwd_launch(hipStream_t s, const DevOp<V>& op, const double* omega, int r0, int r1, int mode,
                      const cvec<V>* in, cvec<V>* p, const cvec<V>* v, cvec<V>* y, const SysState* st,
                      cvec<V>* zl) {
    const dim3 g(sweep_grid<NS, SUB>(op, r1 - r0));
    const size_t lds = LANES ? lanes_lds<NS>(op) : 0;
    if (mode == kInPUpdate)
        hipLaunchKernelGGL((k_sweep_fwd<NS, V, SUB, kInPUpdate, LANES>), g, dim3(kBlock), lds, s, op, omega, r0,
                           r1, in, p, v, y, st, zl);
    else if (mode == kInSUpdate)
        hipLaunchKernelGGL((k_sweep_fwd<NS, V, SUB, kInSUpdate, LANES>), g, dim3(kBlock), lds, s, op, omega, r0,
                           r1, in, p, v, y, st, zl);
    else
        hipLaunchKernelGGL((k_sweep_fwd<NS, V, SUB, kInPlain, LANES>), g, dim3(kBlock), lds, s, op, omega, r0,
                           r1, in, p, v, y, st, zl);
}
template <int NS, typename V, int SUB>
void sweep_fwd_at(hipStream_t s, const DevOp<V>& op, const double* omega, int r0, int r1, int mode,
                  const cvec<V>* in, cvec<V>* p, const cvec<V>* v, cvec<V>* y, const SysState* st,
                  cvec<V>* zl) {
    if constexpr (SUB == 1 && wave_gather_width<NS>()) {
        if (lanes_on<NS>(op)) {
            sweep_fwd_launch<NS, V, 1, true>(s, op, omega, r0, r1, mode, in, p, v, y, st, zl);
            return;
        }
    }
    sweep_fwd_launch<NS, V, SUB, false>(s, op, omega, r0, r1, mode, in, p, v, y, st, zl);
}
template <int NS, typename V, int SUB>
void sweep_bwd_at(hipStream_t s, const DevOp<V>& op, const double* omega, int r0, int r1,
                  const cvec<V>* y, cvec<V>* z, const SysState* st) {
    const dim3 g(sweep_grid<NS, SUB>(op, r1 - r0));
    if constexpr (SUB == 1 && wave_gather_width<NS>()) {
        if (lanes_on<NS>(op)) {
            hipLaunchKernelGGL((k_sweep_bwd<NS, V, 1, true>), g, dim3(kBlock), lanes_lds<NS>(op), s, op, omega,
                               r0, r1, y, z, st);
            return;
        }
    }
    hipLaunchKernelGGL((k_sweep_bwd<NS, V, SUB, false>), g, dim3(kBlock), 0, s, op, omega, r0, r1, y, z, st);
}

template <int NS, typename V>
void Launch<NS, V>::sweep_fwd(hipStream_t s, const DevOp<V>& op, const double* omega, int r0,
                              int r1, int mode, const cvec<V>* in, cvec<V>* p, const cvec<V>* v,
                              cvec<V>* y, const SysState* st, cvec<V>* zl) {
    const int sub = op.shape[width_index(NS)].sweep_sub;
    if (sub == 8 && sub_ok<NS, 8>()) sweep_fwd_at<NS, V, (sub_ok<NS, 8>() ? 8 : 1)>(s, op, omega, r0, r1, mode, in, p, v, y, st, zl);
    else if (sub == 4 && sub_ok<NS, 4>()) sweep_fwd_at<NS, V, (sub_ok<NS, 4>() ? 4 : 1)>(s, op, omega, r0, r1, mode, in, p, v, y, st, zl);
    else if (sub == 2 && sub_ok<NS, 2>()) sweep_fwd_at<NS, V, (sub_ok<NS, 2>() ? 2 : 1)>(s, op, omega, r0, r1, mode, in, p, v, y, st, zl);
    else sweep_fwd_at<NS, V, 1>(s, op, omega, r0, r1, mode, in, p, v, y, st, zl);
}
template <int NS, typename V>
void Launch<NS, V>::sweep_bwd(hipStream_t s, const DevOp<V>& op, const double* omega, int r0,
                              int r1, const cvec<V>* y, cvec<V>* z, const SysState* st) {
    const int sub = op.shape[width_index(NS)].sweep_sub;
    if (sub == 8 && sub_ok<NS, 8>()) sweep_bwd_at<NS, V, (sub_ok<NS, 8>() ? 8 : 1)>(s, op, omega, r0, r1, y, z, st);
    else if (sub == 4 && sub_ok<NS, 4>()) sweep_bwd_at<NS, V, (sub_ok<NS, 4>() ? 4 : 1)>(s, op, omega, r0, r1, y, z, st);
    else if (sub == 2 && sub_ok<NS, 2>()) sweep_bwd_at<NS, V, (sub_ok<NS, 2>() ? 2 : 1)>(s, op, omega, r0, r1, y, z, st);
    else sweep_bwd_at<NS, V, 1>(s, op, omega, r0, r1, y, z, st);
}
template <int NS, typename V>
void Launch<NS, V>::spmv_v(hipStream_t s, const DevOp<V>& op, const double* omega,
                           const cvec<V>* z, const cvec<V>* ys, const cvec<V>* r0, cvec<V>* v,
                           SysState* st, double* part) {
    const int g = row_grid<NS>(op.n, op.shape[width_index(NS)].row_cap);
    if (lanes_on<NS>(op))
        hipLaunchKernelGGL((k_spmv_v<NS, V, wave_gather_width<NS>()>), dim3(g), dim3(kBlock), lanes_lds<NS>(op), s,
                           op, omega, z, ys, r0, v, st, part);
    else
        hipLaunchKernelGGL((k_spmv_v<NS, V, false>), dim3(g), dim3(kBlock), 0, s, op, omega, z, ys, r0, v, st,
                           part);
    hipLaunchKernelGGL((k_spmv_v_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
template <int NS, typename V>
void Launch<NS, V>::spmv_t(hipStream_t s, const DevOp<V>& op, const double* omega,
                           const cvec<V>* z, const cvec<V>* ys, const cvec<V>* sv, cvec<V>* t,
                           SysState* st, double* part) {
    const int g = row_grid<NS>(op.n, op.shape[width_index(NS)].row_cap);
    if (lanes_on<NS>(op))
        hipLaunchKernelGGL((k_spmv_t<NS, V, wave_gather_width<NS>()>), dim3(g), dim3(kBlock), lanes_lds<NS>(op), s,
                           op, omega, z, ys, sv, t, st, part);
    else
        hipLaunchKernelGGL((k_spmv_t<NS, V, false>), dim3(g), dim3(kBlock), 0, s, op, omega, z, ys, sv, t, st,
                           part);
    hipLaunchKernelGGL((k_spmv_t_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
template <int NS, typename V>
void Launch<NS, V>::bicg_xr(hipStream_t s, int n, cvec<V>* x, const cvec<V>* ph, const cvec<V>* sh,
                            const cvec<V>* sv, const cvec<V>* t, const cvec<V>* r0, cvec<V>* r,
                            SysState* st, double* part) {
    const int g = row_grid<NS>(n);
    hipLaunchKernelGGL((k_bicg_xr<NS, V>), dim3(g), dim3(kBlock), 0, s, n, x, ph, sh, sv, t, r0, r, st,
                       part);
    hipLaunchKernelGGL((k_bicg_xr_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
template <int NS, typename V>
void Launch<NS, V>::spmv(hipStream_t s, const DevOp<V>& op, const double* omega, const cvec<V>* x,
                         const cvec<V>* b, cvec<V>* y) {
    if constexpr (wave_gather_width<NS>()) {
        const DevTiles& T = op.tiles[width_index(NS)];
        if (T.ntiles > 0 && op.vtab && op.shape[width_index(NS)].tiled) {
            constexpr int W = kXcds * kTileGroup;
            const size_t lds = tile_table_bytes<V>(op.ntab) + size_t(T.max_stage) * NS * sizeof(cvec<V>);
            hipLaunchKernelGGL((k_spmv_tiled<NS, V>), dim3((T.ntiles + W - 1) / W * W), dim3(kTileBlock), lds, s,
                               op, omega, x,
                               b, y);
            return;
        }
    }
    if (lanes_on<NS>(op)) {
        hipLaunchKernelGGL((k_spmv<NS, V, wave_gather_width<NS>()>), dim3(row_grid<NS>(op.n, kStreamCap)),
                           dim3(kBlock), lanes_lds<NS>(op), s, op, omega, x, b, y);
    } else {
        hipLaunchKernelGGL((k_spmv<NS, V, false>), dim3(row_grid<NS>(op.n, kStreamCap)), dim3(kBlock), 0, s,
                           op, omega, x, b, y);
    }
}
template <int NS, typename V>
void Launch<NS, V>::fill_zero(hipStream_t s, size_t count, cvec<V>* x) {
    hipLaunchKernelGGL((k_fill_zero<NS, V>), dim3(grid_for(count, kStreamCap)), dim3(kBlock), 0, s,
                       count, x);
}
template <int NS, typename V>
void Launch<NS, V>::interp_hanging(hipStream_t s, int npairs, const int* tri, cvec<V>* x) {
    if (npairs <= 0) return;
    hipLaunchKernelGGL((k_interp_hanging<NS, V>), dim3(grid_for(size_t(npairs) * NS, kStreamCap)),
                       dim3(kBlock), 0, s, npairs, tri, x);
}

template <int NS>
void Launch64<NS>::true_res(hipStream_t s, const DevOp<double>& op, const double* omega,
                            const double2* x, const double2* b, double2* r, SysState* st,
                            double* part) {
    const int g = row_grid<NS>(op.n);
    hipLaunchKernelGGL((k_true_res<NS>), dim3(g), dim3(kBlock), 0, s, op, omega, x, b, r, st, part);
    hipLaunchKernelGGL((k_true_res_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
template <int NS>
void Launch64<NS>::scale_guess(hipStream_t s, const DevOp<double>& op, const double* omega,
                               const double2* b, double2* x, SysState* st, double* part) {
    const int g = row_grid<NS>(op.n);
    hipLaunchKernelGGL((k_guess_dots<NS>), dim3(g), dim3(kBlock), 0, s, op, omega, x, b, part);
    hipLaunchKernelGGL((k_guess_dots_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
    hipLaunchKernelGGL((k_scale_guess<NS>), dim3(grid_for(size_t(op.n) * NS, kStreamCap)), dim3(kBlock),
                       0, s, op.n, st, x);
}
template <int NS>
void Launch64<NS>::zero_systems(hipStream_t s, int npad, const int* flags, double2* x) {
    hipLaunchKernelGGL((k_zero_systems<NS>), dim3(grid_for(size_t(npad) * NS, kStreamCap)),
                       dim3(kBlock), 0, s, npad, flags, x);
}
template <int NS>
void Launch64<NS>::copy_systems(hipStream_t s, int npad, const int* flags, const double2* src, double2* dst) {
    hipLaunchKernelGGL((k_copy_systems<NS>), dim3(grid_for(size_t(npad) * NS, kStreamCap)), dim3(kBlock), 0, s,
                       npad, flags, src, dst);
}
template <int NS>
void Launch64<NS>::forward_total(hipStream_t s, int npad, const unsigned char* bc,
                                 const double2* epri, const double2* b, double2* x) {
    hipLaunchKernelGGL((k_forward_total<NS>), dim3(grid_for(size_t(npad) * NS, kStreamCap)),
                       dim3(kBlock), 0, s, npad, bc, epri, b, x);
}
template <int NS>
void Launch64<NS>::fix_constrained(hipStream_t s, int npad, const unsigned char* cons, const double2* b,
                                   double2* x) {
    hipLaunchKernelGGL((k_fix_constrained<NS>), dim3(grid_for(size_t(npad) * NS, kStreamCap)),
                       dim3(kBlock), 0, s, npad, cons, b, x);
}
template <int NS>
void Launch64<NS>::add_correction(hipStream_t s, size_t count, const float2* d, double2* x) {
    hipLaunchKernelGGL((k_add_correction<NS>), dim3(grid_for(count, kStreamCap)), dim3(kBlock), 0,
                       s, count, d, x);
}
template <int NS>
void Launch64<NS>::demote(hipStream_t s, size_t count, const double2* src, const double* scale,
                          float2* dst) {
    hipLaunchKernelGGL((k_demote<NS>), dim3(grid_for(count, kStreamCap)), dim3(kBlock), 0, s, count,
                       src, scale, dst);
}
template <int NS>
void Launch64<NS>::add_scaled(hipStream_t s, size_t count, const float2* d, const double* scale,
                              double2* x) {
    hipLaunchKernelGGL((k_add_scaled<NS>), dim3(grid_for(count, kStreamCap)), dim3(kBlock), 0, s,
                       count, d, scale, x);
}
template <int NS>
void Launch64<NS>::phi0(hipStream_t s, int nn, const int* nt_ptr, const int* nt_edge,
                        const double* nt_val, const unsigned char* dead, const double* omega,
                        const double2* b, double2* phi0) {
    hipLaunchKernelGGL((k_phi0<NS>), dim3(grid_for(size_t(nn) * NS, kStreamCap)), dim3(kBlock), 0, s,
                       nn, nt_ptr, nt_edge, nt_val, dead, omega, b, phi0);
}
template <int NS, typename V>
void LaunchDc<NS, V>::divj(hipStream_t s, int nn, int it, double tol, int cg_max, double pcg_tol,
                           const int* nt_ptr, const int* nt_edge, const double* nt_val, const double* ms,
                           const unsigned char* bc_node, const V* ldiag, const double2* E, const double2* phi0,
                           C* r, C* z, C* p, C* phi, SysState* st, double* part) {
    const int g = row_grid<NS>(nn);
    hipLaunchKernelGGL((k_dc_divj<NS, V>), dim3(g), dim3(kBlock), 0, s, nn, nt_ptr, nt_edge, nt_val, ms,
                       bc_node, ldiag, E, phi0, r, z, p, phi, st, part);
    hipLaunchKernelGGL((k_dc_divj_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, it, tol, cg_max, pcg_tol, st, part, g);
}
template <int NS, typename V>
void LaunchDc<NS, V>::lp(hipStream_t s, int nn, const int* lptr, const int* lcol, const V* lval, const C* p,
                         C* ap, SysState* st, double* part) {
    const int g = row_grid<NS>(nn);
    // Entries shared among a row's lanes from 8 systems (bitwise the same
    // sums; the 16-system FP64 product on a 483 K-node Laplacian 120 -> 108 us).
    if constexpr (NS >= 8)
        hipLaunchKernelGGL((k_dc_lp<NS, true, V>), dim3(g), dim3(kBlock), 0, s, nn, lptr, lcol, lval, p, ap, st,
                           part);
    else
        hipLaunchKernelGGL((k_dc_lp<NS, false, V>), dim3(g), dim3(kBlock), 0, s, nn, lptr, lcol, lval, p, ap, st,
                           part);
    hipLaunchKernelGGL((k_dc_lp_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
template <int NS, typename V>
void LaunchDc<NS, V>::update(hipStream_t s, int nn, const V* ldiag, const C* p, const C* ap, C* phi, C* r,
                             SysState* st, double* part) {
    const int g = row_grid<NS>(nn);
    hipLaunchKernelGGL((k_dc_update<NS, V>), dim3(g), dim3(kBlock), 0, s, nn, ldiag, p, ap, phi, r, st, part);
    hipLaunchKernelGGL((k_dc_update_fin<NS>), dim3(1), dim3(kFinBlock), 0, s, st, part, g);
}
template <int NS, typename V>
void LaunchDc<NS, V>::dir(hipStream_t s, int nn, const V* ldiag, const C* r, C* p, const SysState* st) {
    hipLaunchKernelGGL((k_dc_p<NS, V>), dim3(grid_for(size_t(nn) * NS, kStreamCap)), dim3(kBlock), 0, s,
                       nn, ldiag, r, p, st);
}
template <int NS, typename V>
void LaunchDc<NS, V>::correct(hipStream_t s, int npad, const int* g_col, const double* g_val,
                              const unsigned char* bc_edge, const C* phi, double2* E, const SysState* st) {
    hipLaunchKernelGGL((k_dc_correct<NS, V>), dim3(grid_for(size_t(npad) * NS, kStreamCap)),
                       dim3(kBlock), 0, s, npad, g_col, g_val, bc_edge, phi, E, st);
}

__global__ void k_to_f32(size_t n, const double* __restrict__ src, float* __restrict__ dst) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = float(src[i]);
}
void launch_to_f32(hipStream_t s, size_t n, const double* src, float* dst) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, kStreamCap)), dim3(kBlock), 0, s, n, src, dst);
}

template <typename V>
void launch_regroup(hipStream_t s, int rows, int na, int nb, const int* map, cvec<V>* wide,
                    cvec<V>* narrow, bool to_narrow) {
    hipLaunchKernelGGL((k_regroup<V>), dim3(grid_for(size_t(rows) * nb, kStreamCap)), dim3(kBlock),
                       0, s, rows, na, nb, map, wide, narrow, to_narrow ? 1 : 0);
}
template void launch_regroup<double>(hipStream_t, int, int, int, const int*, double2*, double2*, bool);
template void launch_regroup<float>(hipStream_t, int, int, int, const int*, float2*, float2*, bool);

void launch_sigma_edges(hipStream_t s, int npad, int nc, const int* perm, const int* adj_ptr,
                        const int* adj_cell, const double* area, const double* len,
                        const unsigned char* massless, const unsigned char* dead,
                        const unsigned char* constrained, const unsigned char* dc_dead,
                        const unsigned char* in_pattern, const double* sigma, double* me, float* me32,
                        double* ms, int* err) {
    hipLaunchKernelGGL(k_sigma_edges, dim3(grid_for(size_t(npad), kStreamCap)), dim3(kBlock), 0, s, npad, nc,
                       perm, adj_ptr, adj_cell, area, len, massless, dead, constrained, dc_dead, in_pattern,
                       sigma, me, me32, ms, err);
}
void launch_sigma_laplacian(hipStream_t s, int nn, const int* nt_ptr, const int* nt_edge, const double* nt_val,
                            const double* ms, const int* lptr, const int* lcol, const int* ledge,
                            const int* g_col, const double* g_val, const unsigned char* bc_node, double* lval,
                            double* ldiag) {
    hipLaunchKernelGGL(k_sigma_laplacian, dim3(grid_for(size_t(nn), kStreamCap)), dim3(kBlock), 0, s, nn, nt_ptr,
                       nt_edge, nt_val, ms, lptr, lcol, ledge, g_col, g_val, bc_node, lval, ldiag);
}

void launch_perturbation_rhs(hipStream_t s, int n, int nsys, int nc, const int* e_ptr, const int* e_cells,
                             const int* adj_ptr, const int* adj_cell, const double* len, const double* sigma,
                             const double* dsig, const double* omega, const double2* ebg, double2* rhs) {
    hipLaunchKernelGGL(k_perturbation_rhs, dim3(grid_for(size_t(n) * nsys, kStreamCap)), dim3(kBlock), 0, s, n,
                       nsys, nc, e_ptr, e_cells, adj_ptr, adj_cell, len, sigma, dsig, omega, ebg, rhs);
}

void launch_bicg_begin(hipStream_t s, SysState* st, unsigned mask, int maxiter, double tol) {
    hipLaunchKernelGGL(k_bicg_begin, dim3(1), dim3(64), 0, s, st, mask, maxiter, tol);
}
void launch_dc_begin(hipStream_t s, SysState* st, unsigned mask) {
    hipLaunchKernelGGL(k_dc_begin, dim3(1), dim3(64), 0, s, st, mask);
}

#define M3D_INST(NS)                     \
    template struct Launch<NS, double>;  \
    template struct Launch<NS, float>;   \
    template struct Launch64<NS>;        \
    template struct LaunchDc<NS, double>; \
    template struct LaunchDc<NS, float>;
M3D_INST(1)
M3D_INST(2)
M3D_INST(4)
M3D_INST(8)
M3D_INST(16)
M3D_INST(32)
#undef M3D_INST

}  // namespace dev
}  // namespace m3d
