// Coarse-grid neural operator for the initial field e0: four 3x3x3 dilated
// conv layers on a 32^3 grid, BF16 storage, FP32 accumulation on the CDNA
// matrix cores (v_mfma_f32_32x32x16_bf16; v_mfma_f32_32x32x8_bf16 over the
// padded 8-channel input of the forward pass), trained with Adam on the
// device.
//
// Every conv pass is an implicit GEMM whose 32x32 MFMA tile is one x-row of
// the grid (32 voxels) against 32 channels:
//   forward      out[v, co]  = sum_{tap, ci} in[v + d*tap, ci] * W[tap][co][ci]
//   input grad   din[v, ci]  = sum_{tap, co} dz[v - d*tap, co] * W[tap][co][ci]
//   weight grad  dW[tap][co][ci] = sum_v dz[v, co] * in[v + d*tap, ci]
// Activations are channels-last [batch][z][y][x][C]; zero padding outside
// the grid.  The layer has no reference counterpart: the reference starts
// every forward solve from zero (forward_solver_fv.cpp:353).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "m3d_net.h"

namespace {

constexpr int G = 32;                 // grid points per axis
constexpr int V = G * G * G;          // voxels per sample
constexpr int NL = 4;                 // conv layers
constexpr int HID = 32;               // hidden channels (and padded output channels)
constexpr int CIN0 = 8;               // input channels after padding (2 used)
constexpr int COUT = 12;              // output channels (2 pol x 3 comp x re/im)
constexpr int CIN_USED = 2;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16 __attribute__((ext_vector_type(16)));

// 32x32x16 BF16 MFMA: lane l holds row (A) / column (B) l % 32 and the eight
// k values 8 (l / 32) .. 8 (l / 32) + 7 of the 16-wide k step; the
// accumulator layout is that of the 32x32x8 form.
__device__ inline f16 mfma16(s8 a, s8 b, f16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
                                                   0);
}

// Taps whose operands are in flight ahead of the MFMAs of the current tap.
constexpr int kTapStages = 3;

__device__ inline unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return static_cast<unsigned short>(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }

__device__ inline void tap_offsets(int tap, int& dz, int& dy, int& dx) {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
}

// Forward conv of one layer.  One wavefront = one x-row (b, z, y) x 32 output
// channels.  Lane l: voxel x = l % 32 for the A operand, output channel
// l % 32 for B; q = l / 32 selects channels 4q..4q+3 of each 8-wide k step
// (CIN = 8) or 8q..8q+7 of each 16-wide one (CIN a multiple of 16).
// Accumulator r of lane l is D[x = 8(r/4) + 4q + r%4][co = l % 32].  With
// 16-wide steps the 27 taps are unrolled and the operands of the next
// kTapStages - 1 taps are loaded while the current tap's MFMAs run.
template <int CIN, bool RELU, bool OUTF32>
__global__ __launch_bounds__(256) void k_conv_fwd(int batch, int dil,
                                                  const unsigned short* __restrict__ in,
                                                  const unsigned short* __restrict__ wt,
                                                  const float* __restrict__ bias,
                                                  unsigned short* __restrict__ out_bf,
                                                  float* __restrict__ out_f, int cout_store) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    if (row >= batch * G * G) return;
    const int b = row / (G * G), z = (row / G) % G, y = row % G;
    const int m = lane & 31, q = lane >> 5;
    f16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    if constexpr (CIN % 16 == 0) {
        constexpr int NK = CIN / 16;
        s8 a[kTapStages][NK], w[kTapStages][NK];
        auto fetch = [&](int tap, s8 (&A)[NK], s8 (&W)[NK]) {
            int dz, dy, dx;
            tap_offsets(tap, dz, dy, dx);
            const int zz = z + dz * dil, yy = y + dy * dil, xx = m + dx * dil;
            const bool ok = zz >= 0 && zz < G && yy >= 0 && yy < G && xx >= 0 && xx < G;
            const size_t vin = ((size_t(b) * G + (ok ? zz : 0)) * G + (ok ? yy : 0)) * G + (ok ? xx : 0);
            const unsigned short* src = in + vin * CIN;
            const unsigned short* wsrc = wt + (size_t(tap) * HID + m) * CIN;
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                s8 t = {0, 0, 0, 0, 0, 0, 0, 0};
                if (ok) t = *reinterpret_cast<const s8*>(src + 16 * k + 8 * q);
                A[k] = t;
                W[k] = *reinterpret_cast<const s8*>(wsrc + 16 * k + 8 * q);
            }
        };
#pragma unroll
        for (int t = 0; t < kTapStages - 1; ++t) fetch(t, a[t], w[t]);
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            if (tap + kTapStages - 1 < 27)
                fetch(tap + kTapStages - 1, a[(tap + kTapStages - 1) % kTapStages],
                      w[(tap + kTapStages - 1) % kTapStages]);
#pragma unroll
            for (int k = 0; k < NK; ++k) acc = mfma16(a[tap % kTapStages][k], w[tap % kTapStages][k], acc);
        }
    } else {
        static_assert(CIN == 8, "8-wide k steps serve the padded input layer");
        s4 a[kTapStages], w[kTapStages];
        auto fetch = [&](int tap, s4& A, s4& W) {
            int dz, dy, dx;
            tap_offsets(tap, dz, dy, dx);
            const int zz = z + dz * dil, yy = y + dy * dil, xx = m + dx * dil;
            const bool ok = zz >= 0 && zz < G && yy >= 0 && yy < G && xx >= 0 && xx < G;
            const size_t vin = ((size_t(b) * G + (ok ? zz : 0)) * G + (ok ? yy : 0)) * G + (ok ? xx : 0);
            s4 t = {0, 0, 0, 0};
            if (ok) t = *reinterpret_cast<const s4*>(in + vin * CIN + 4 * q);
            A = t;
            W = *reinterpret_cast<const s4*>(wt + (size_t(tap) * HID + m) * CIN + 4 * q);
        };
#pragma unroll
        for (int t = 0; t < kTapStages - 1; ++t) fetch(t, a[t], w[t]);
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            if (tap + kTapStages - 1 < 27)
                fetch(tap + kTapStages - 1, a[(tap + kTapStages - 1) % kTapStages],
                      w[(tap + kTapStages - 1) % kTapStages]);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[tap % kTapStages], w[tap % kTapStages], acc, 0, 0, 0);
        }
    }
    const int co = m;
    const float bc = bias[co];
    const size_t vrow = ((size_t(b) * G + z) * G + y) * G;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int x = 8 * (r >> 2) + 4 * q + (r & 3);
        float v = acc[r] + bc;
        if (RELU) v = v > 0.f ? v : 0.f;
        if (OUTF32) {
            if (co < cout_store) out_f[(vrow + x) * cout_store + co] = v;
        } else {
            out_bf[(vrow + x) * HID + co] = f2bf(v);
        }
    }
}

// Forward conv of a 32-channel layer with the layer's weights in LDS: a
// workgroup of kConvWaves wavefronts copies W (27 taps x 32 x 32 BF16, 54 KB)
// into LDS once and its wavefronts then walk x-rows as k_conv_fwd does,
// reading the B operands from LDS instead of each wavefront reading all 54 KB
// from the caches.  16-byte units are stored XOR-swizzled within their
// 64-byte (tap, co) row, unit c at c ^ swz(co), so the 16 lanes that read one
// unit of 16 consecutive co rows hit distinct banks.
constexpr int kConvWaves = 8;
constexpr int kConvWBytes = 27 * HID * HID * 2;

__device__ inline int wswz(int r) { return (((r >> 1) ^ (r >> 3)) & 1) | ((r >> 1) & 2); }

__device__ inline void stage_weights(const unsigned short* __restrict__ w, s8* __restrict__ lds) {
    constexpr int units = kConvWBytes / 16;  // 4 units per 32-channel row
    for (int i = threadIdx.x; i < units; i += blockDim.x) {
        const int r = i >> 2, c = i & 3;
        lds[r * 4 + (c ^ wswz(r & 31))] = reinterpret_cast<const s8*>(w)[i];
    }
}

__device__ inline s8 lds_unit(const s8* __restrict__ lds, int tap, int r, int c) {
    return lds[(tap * HID + r) * 4 + (c ^ wswz(r))];
}

// BWD: the input gradient din[v, ci] = sum_{tap, co} dz[v - d*tap, co] *
// W[tap][co][ci] of a 32-channel layer, from the transposed weights
// wb[tap][ci][co] and an already masked dz, stored as BF16 without bias.
template <bool BWD, bool RELU, bool OUTF32>
__global__ __launch_bounds__(kConvWaves * 64) void k_conv_fwd_w(int batch, int dil,
                                                                const unsigned short* __restrict__ in,
                                                                const unsigned short* __restrict__ wt,
                                                                const float* __restrict__ bias,
                                                                unsigned short* __restrict__ out_bf,
                                                                float* __restrict__ out_f, int cout_store) {
    extern __shared__ s8 wl[];
    stage_weights(wt, wl);
    __syncthreads();
    constexpr int NK = HID / 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 31, q = lane >> 5;
    const int rows = batch * G * G;
    const float bc = BWD ? 0.f : bias[m];
    const int sg = BWD ? -1 : 1;
    for (int row = blockIdx.x * kConvWaves + wave; row < rows; row += gridDim.x * kConvWaves) {
        const int b = row / (G * G), z = (row / G) % G, y = row % G;
        f16 acc;
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
        s8 a[kTapStages][NK];
        auto fetch = [&](int tap, s8 (&A)[NK]) {
            int dz, dy, dx;
            tap_offsets(tap, dz, dy, dx);
            const int zz = z + sg * dz * dil, yy = y + sg * dy * dil, xx = m + sg * dx * dil;
            const bool ok = zz >= 0 && zz < G && yy >= 0 && yy < G && xx >= 0 && xx < G;
            const size_t vin = ((size_t(b) * G + (ok ? zz : 0)) * G + (ok ? yy : 0)) * G + (ok ? xx : 0);
            const unsigned short* src = in + vin * HID;
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                s8 t = {0, 0, 0, 0, 0, 0, 0, 0};
                if (ok) t = *reinterpret_cast<const s8*>(src + 16 * k + 8 * q);
                A[k] = t;
            }
        };
        // B operands one tap ahead from LDS; a scheduling barrier per tap
        // keeps the compiler from hoisting every tap's LDS reads (which
        // would take 216 registers).
        s8 wv[2][NK];
#pragma unroll
        for (int t = 0; t < kTapStages - 1; ++t) fetch(t, a[t]);
#pragma unroll
        for (int k = 0; k < NK; ++k) wv[0][k] = lds_unit(wl, 0, m, 2 * k + q);
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            if (tap + kTapStages - 1 < 27) fetch(tap + kTapStages - 1, a[(tap + kTapStages - 1) % kTapStages]);
            if (tap + 1 < 27) {
#pragma unroll
                for (int k = 0; k < NK; ++k) wv[(tap + 1) & 1][k] = lds_unit(wl, tap + 1, m, 2 * k + q);
            }
#pragma unroll
            for (int k = 0; k < NK; ++k) acc = mfma16(a[tap % kTapStages][k], wv[tap & 1][k], acc);
            __builtin_amdgcn_sched_barrier(0);
        }
        const size_t vrow = ((size_t(b) * G + z) * G + y) * G;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int x = 8 * (r >> 2) + 4 * q + (r & 3);
            float v = acc[r] + bc;
            if (RELU) v = v > 0.f ? v : 0.f;
            if (OUTF32) {
                if (m < cout_store) out_f[(vrow + x) * cout_store + m] = v;
            } else {
                out_bf[(vrow + x) * HID + m] = f2bf(v);
            }
        }
    }
}

// Weight gradient partials: block (chunk, dz dy) sums dz[v, co] * in[v + d*tap, ci]
// over the rows of its chunk for the three taps (dz, dy, dx = -1, 0, 1);
// M = co (32), N = ci (CIN, zero-padded to 32), K = voxels.  The
// contraction runs over voxels, the strided dimension of channels-last
// activations, so each wavefront stages one x-row of dz (already masked by
// the ReLU) and the row (z + dz d, y + dy d) of the input, shifted by dx d
// for each of the three taps, through LDS: 16-byte global loads, written
// transposed into wave-private tiles [channel][x] (row stride kTileStride
// keeps every 8-voxel operand 16-byte aligned), then read back as the
// operands of 16-deep MFMAs; the dz operands are read once for the three
// taps.  A tile belongs to one wavefront and LDS serves a wavefront's
// accesses in order, so no block barrier is needed between rows (the
// wavefronts of a block may take different row counts).
constexpr int kTileStride = 40;

// Tile position of voxel x (even) in channel row r: the four 16-byte units
// of a row (8 voxels each) are XOR-swizzled by (r / 8) % 4, so the 4-byte
// writes of the four channel octets a wavefront transposes at once fall on
// different banks (a row stride of 40 shorts alone maps them onto the same
// ones), while every 8-voxel operand stays one aligned 16-byte unit.
__device__ inline int tile_at(int r, int x) {
    return r * kTileStride + 8 * ((x >> 3) ^ ((r >> 3) & 3)) + (x & 7);
}

template <int CIN>
__global__ __launch_bounds__(256) void k_conv_bwd_w(int batch, int dil, int rows_per_chunk,
                                                    const unsigned short* __restrict__ in,
                                                    const unsigned short* __restrict__ dout,
                                                    float* __restrict__ part) {
    // Per wavefront a dz tile and three input tiles; after the rows the same
    // LDS holds the wavefronts' accumulators for the block sum (40 KB, three
    // blocks per CU).
    constexpr int kT = HID * kTileStride;
    __shared__ __attribute__((aligned(16))) unsigned short lds[4 * 4 * kT];
    auto tdz = reinterpret_cast<unsigned short(*)[kT]>(lds);
    auto tin = reinterpret_cast<unsigned short(*)[3][kT]>(lds + 4 * kT);
    auto red = reinterpret_cast<float(*)[16][64]>(lds);
    static_assert(4 * 16 * 64 * sizeof(float) <= sizeof(lds), "the block sum fits the tiles' LDS");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int chunk = blockIdx.x, dzy = blockIdx.y;  // dzy = 3 (dz + 1) + (dy + 1)
    const int m = lane & 31, q = lane >> 5;
    const int dz = dzy / 3 - 1, dy = dzy % 3 - 1;
    unsigned short* Tdz = tdz[wave];
    // Input channels CIN..31 of the B tiles stay zero.
    for (int k = lane; k < 3 * kT; k += 64) tin[wave][k / kT][k % kT] = 0;
    f16 acc[3];
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < 16; ++i) acc[d][i] = 0.f;
    const int rows = batch * G * G;
    const int r0 = chunk * rows_per_chunk, r1 = min(rows, r0 + rows_per_chunk);
    // Global loads of a row: lane l takes voxels x, x + 1 (x = 2 (l / 4)) of
    // dz channels c0 .. c0 + 7 (c0 = 8 (l % 4)); lanes below NP take a voxel
    // pair and channel octet of the input row for each of the three shifts.
    // The next row's loads are issued before this row's tiles are read.
    constexpr int NO = CIN / 8;  // channel octets per voxel
    constexpr int NP = 16 * NO;  // (voxel pair, octet) items of the input
    const int xg = 2 * (lane >> 2), cg = (lane & 3) * 8;
    const int xi = 2 * (lane / NO), ci0 = (lane % NO) * 8;
    s8 gdz[2], gin[3][2];
    auto load = [&](int row) {
        const int b = row / (G * G), z = (row / G) % G, y = row % G;
        const size_t vrow = ((size_t(b) * G + z) * G + y) * G;
        const int zz = z + dz * dil, yy = y + dy * dil;
        const bool okzy = zz >= 0 && zz < G && yy >= 0 && yy < G;
        const size_t vsrc_row = ((size_t(b) * G + (okzy ? zz : 0)) * G + (okzy ? yy : 0)) * G;
#pragma unroll
        for (int h = 0; h < 2; ++h) gdz[h] = *reinterpret_cast<const s8*>(dout + (vrow + xg + h) * HID + cg);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int xs = xi + h + (d - 1) * dil;
                s8 v = {0, 0, 0, 0, 0, 0, 0, 0};
                if (lane < NP && okzy && xs >= 0 && xs < G)
                    v = *reinterpret_cast<const s8*>(in + (vsrc_row + xs) * CIN + ci0);
                gin[d][h] = v;
            }
    };
    if (r0 + wave < r1) load(r0 + wave);
    for (int row = r0 + wave; row < r1; row += 4) {
        __builtin_amdgcn_wave_barrier();
        // Transposed tiles: each channel's voxel pair is one 4-byte word.
#pragma unroll
        for (int j = 0; j < 8; ++j)
            *reinterpret_cast<unsigned*>(Tdz + tile_at(cg + j, xg)) =
                unsigned(static_cast<unsigned short>(gdz[0][j])) |
                (unsigned(static_cast<unsigned short>(gdz[1][j])) << 16);
        if (lane < NP) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                unsigned short* Tin = tin[wave][d];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    *reinterpret_cast<unsigned*>(Tin + tile_at(ci0 + j, xi)) =
                        unsigned(static_cast<unsigned short>(gin[d][0][j])) |
                        (unsigned(static_cast<unsigned short>(gin[d][1][j])) << 16);
            }
        }
        if (row + 4 < r1) load(row + 4);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int s16 = 0; s16 < 2; ++s16) {
            const s8 a = *reinterpret_cast<const s8*>(Tdz + tile_at(m, 16 * s16 + 8 * q));
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const s8 w = *reinterpret_cast<const s8*>(tin[wave][d] + tile_at(m, 16 * s16 + 8 * q));
                acc[d] = mfma16(a, w, acc[d]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    }
    __syncthreads();  // every wavefront is done with its tiles
    for (int d = 0; d < 3; ++d) {
        for (int r = 0; r < 16; ++r) red[wave][r][lane] = acc[d][r];
        __syncthreads();
        if (wave == 0) {
            const int tap = dzy * 3 + d;
            for (int r = 0; r < 16; ++r) {
                const float v = red[0][r][lane] + red[1][r][lane] + red[2][r][lane] + red[3][r][lane];
                const int co = 8 * (r >> 2) + 4 * q + (r & 3);  // D row = M index = co
                const int ci = m;                               // D column = N index = ci
                if (ci < CIN) part[((size_t(chunk) * 27 + tap) * HID + co) * CIN + ci] = v;
            }
        }
        __syncthreads();
    }
}

// out[i] = sum_c part[c][i] over nchunk partial sets of count values.  A
// 256-thread block takes 32 consecutive outputs; its eight groups of 32
// threads each add every eighth partial set, and the groups' sums are added
// in LDS.
__global__ __launch_bounds__(256) void k_sum_chunks(int nchunk, int count, const float* __restrict__ part,
                                                    float* __restrict__ out) {
    __shared__ float sm[8][32];
    const int j = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int i = blockIdx.x * 32 + j;
    float s = 0.f;
    if (i < count)
        for (int c = g; c < nchunk; c += 8) s += part[size_t(c) * count + i];
    sm[g][j] = s;
    __syncthreads();
    if (g == 0 && i < count) {
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) t += sm[k][j];
        out[i] = t;
    }
}

// Bias gradient partials: per block the sum of dz (masked by act > 0) over
// its voxels, 32 channels.  Thread t loads channels 8 (t % 4) .. + 7 of one
// voxel (16-byte loads), 64 voxels per block step; the 64 voxel lanes are
// added in LDS.
constexpr int kBiasBlocks = 512;
__global__ __launch_bounds__(256) void k_bias_grad(int nvox, const unsigned short* __restrict__ dout,
                                                   const unsigned short* __restrict__ act, float* __restrict__ part) {
    __shared__ float sm[64][33];
    const int o = threadIdx.x & 3, vl = threadIdx.x >> 2;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int v = blockIdx.x * 64 + vl; v < nvox; v += gridDim.x * 64) {
        const s8 g = *reinterpret_cast<const s8*>(dout + size_t(v) * HID + 8 * o);
        s8 a = {1, 1, 1, 1, 1, 1, 1, 1};
        if (act) a = *reinterpret_cast<const s8*>(act + size_t(v) * HID + 8 * o);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (!act || bf2f(static_cast<unsigned short>(a[j])) > 0.f) acc[j] += bf2f(static_cast<unsigned short>(g[j]));
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) sm[vl][8 * o + j] = acc[j];
    __syncthreads();
    // 32 channels x 8 groups of 8 voxel lanes, then the 8 group sums.
    const int c = threadIdx.x & 31, gq = threadIdx.x >> 5;
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) t += sm[gq * 8 + k][c];
    __syncthreads();
    sm[gq][c] = t;
    __syncthreads();
    if (threadIdx.x < 32) {
        float u = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) u += sm[k][threadIdx.x];
        part[blockIdx.x * 32 + threadIdx.x] = u;
    }
}

// out[c] = sum over nb partials of 32 channels: one block of 1024 threads,
// 32 groups each adding every 32nd partial, then the groups in LDS.
__global__ __launch_bounds__(1024) void k_sum_bias(int nb, const float* __restrict__ part, float* __restrict__ out) {
    __shared__ float sm[32][33];
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    float s = 0.f;
    for (int i = g; i < nb; i += 32) s += part[i * 32 + c];
    sm[g][c] = s;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.f;
        for (int k = 0; k < 32; ++k) t += sm[k][threadIdx.x];
        out[threadIdx.x] = t;
    }
}

// dz = dA * 1[act > 0] in place (BF16, 8 channels per thread).
__global__ void k_mask_relu(size_t units, unsigned short* __restrict__ g, const unsigned short* __restrict__ act) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < units; i += size_t(gridDim.x) * blockDim.x) {
        s8 v = reinterpret_cast<s8*>(g)[i];
        const s8 a = reinterpret_cast<const s8*>(act)[i];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (bf2f(static_cast<unsigned short>(a[j])) <= 0.f) v[j] = 0;
        reinterpret_cast<s8*>(g)[i] = v;
    }
}

// Input features [n][2] FP32 -> [n][8] BF16 (zero padded).
__global__ void k_prep_input(int nvox, const float* __restrict__ x, unsigned short* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * CIN0) return;
    const int v = i / CIN0, c = i % CIN0;
    out[i] = c < CIN_USED ? f2bf(x[size_t(v) * CIN_USED + c]) : 0;
}

// MSE loss partials and dL/dout = 2 (out - tgt) / count as BF16 [n][32].
__global__ void k_loss(int nvox, const float* __restrict__ out, const float* __restrict__ tgt,
                       float scale, unsigned short* __restrict__ dout, float* __restrict__ part) {
    __shared__ float sm[256];
    float s = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nvox * HID; i += gridDim.x * blockDim.x) {
        const int v = i / HID, c = i % HID;
        float g = 0.f;
        if (c < COUT) {
            const float d = out[size_t(v) * COUT + c] - tgt[size_t(v) * COUT + c];
            s += d * d;
            g = 2.f * d * scale;
        }
        dout[i] = f2bf(g);
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Every parameter of the four layers in one index space: the weight arrays
// of layers 0-3, then their bias arrays (off[k] .. off[k + 1] for array k).
struct ParamSet {
    float* p[8];
    const float* g[8];
    float* m[8];
    float* v[8];
    unsigned short* wt[4];  // BF16 weights [tap][co][ci]
    unsigned short* wb[4];  // BF16 transposed weights [tap][ci][co] (32-channel inputs)
    int cin[4];
    int off[9];
};
__device__ inline int param_array(const ParamSet& P, int i) {
    int k = 0;
    while (k < 7 && i >= P.off[k + 1]) ++k;
    return k;
}

// Global gradient norm, first stage: kSqBlocks block partials of the sum of
// squares over the parameter index space (fixed slots, fixed order).
constexpr int kSqBlocks = 64;
__global__ __launch_bounds__(256) void k_grad_sq(ParamSet P, float* __restrict__ part) {
    __shared__ float sm[256];
    float s = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < P.off[8]; i += kSqBlocks * 256) {
        const int k = param_array(P, i);
        const float g = P.g[k][i - P.off[k]];
        s += g * g;
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Adam on the FP32 master parameters of every layer in one launch, the
// gradient scaled by min(1, clip / ||g||) (1 when clip <= 0; every block
// adds the kSqBlocks partials in the same order), and the BF16 copies of the
// updated weights that the conv kernels read.
__global__ __launch_bounds__(256) void k_adam_all(ParamSet P, float lr, float b1, float b2, float eps, float c1,
                                                  float c2, float clip, const float* __restrict__ sq) {
    static_assert(kSqBlocks == 64, "one wavefront adds the partials");
    __shared__ float sc;
    if (threadIdx.x < 64) {
        float t = sq[threadIdx.x];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (threadIdx.x == 0) {
            const float nrm = sqrtf(t);
            sc = (clip > 0.f && nrm > clip) ? clip / nrm : 1.f;
        }
    }
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P.off[8]) return;
    const int k = param_array(P, i), j = i - P.off[k];
    const float gi = P.g[k][j] * sc;
    const float m = b1 * P.m[k][j] + (1.f - b1) * gi;
    const float v = b2 * P.v[k][j] + (1.f - b2) * gi * gi;
    P.m[k][j] = m;
    P.v[k][j] = v;
    const float pn = P.p[k][j] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P.p[k][j] = pn;
    if (k < 4) {
        const int cin = P.cin[k];
        const int tap = j / (HID * cin), co = (j / cin) % HID, ci = j % cin;
        const unsigned short h = f2bf(pn);
        P.wt[k][j] = h;
        if (cin == HID) P.wb[k][(size_t(tap) * HID + ci) * HID + co] = h;
    }
}

__global__ void k_weights_bf16(int cin, const float* __restrict__ w, unsigned short* __restrict__ wt,
                               unsigned short* __restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 27 * HID * cin) return;
    const int tap = i / (HID * cin), co = (i / cin) % HID, ci = i % cin;
    const unsigned short h = f2bf(w[i]);
    wt[i] = h;
    if (wb && cin == HID) wb[(size_t(tap) * HID + ci) * HID + co] = h;
}

// Trilinear interpolation of the voxel field (voxel centres of the box
// [x0, x0 + 32 hx] x ...) to edge midpoints: system 2b + pol of the output
// takes channel pol*6 + 2*dir + {re, im} of sample b; edges outside the
// centre lattice get 0.
__global__ void k_prolong(int batch, int n, const float* __restrict__ field,
                          const float* __restrict__ exyz, const unsigned char* __restrict__ edir,
                          float x0, float y0, float z0, float hx, float hy, float hz,
                          double2* __restrict__ out) {
    const size_t total = size_t(batch) * n;
    for (size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x; t < total;
         t += size_t(gridDim.x) * blockDim.x) {
        const int b = int(t / n), e = int(t % n);
        const float fx = (exyz[3 * size_t(e)] - x0) / hx - 0.5f;
        const float fy = (exyz[3 * size_t(e) + 1] - y0) / hy - 0.5f;
        const float fz = (exyz[3 * size_t(e) + 2] - z0) / hz - 0.5f;
        double2 v1 = make_double2(0.0, 0.0), v2 = make_double2(0.0, 0.0);
        if (fx >= 0.f && fy >= 0.f && fz >= 0.f && fx <= G - 1 && fy <= G - 1 && fz <= G - 1) {
            const int ix = min(int(fx), G - 2), iy = min(int(fy), G - 2), iz = min(int(fz), G - 2);
            const float tx = fx - ix, ty = fy - iy, tz = fz - iz;
            const int d = edir[e];
            float a[4] = {0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 8; ++c) {
                const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                const float wgt = (cx ? tx : 1.f - tx) * (cy ? ty : 1.f - ty) * (cz ? tz : 1.f - tz);
                const size_t vox = ((size_t(b) * G + iz + cz) * G + iy + cy) * G + ix + cx;
                const float* f = field + vox * COUT;
                a[0] += wgt * f[2 * d];
                a[1] += wgt * f[2 * d + 1];
                a[2] += wgt * f[6 + 2 * d];
                a[3] += wgt * f[6 + 2 * d + 1];
            }
            v1 = make_double2(a[0], a[1]);
            v2 = make_double2(a[2], a[3]);
        }
        out[(2 * size_t(b)) * n + e] = v1;
        out[(2 * size_t(b) + 1) * n + e] = v2;
    }
}

void chk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { if (p) (void)hipFree(p); }
    void alloc(size_t c) {
        if (p) (void)hipFree(p);
        p = nullptr;
        chk(hipMalloc(&p, std::max<size_t>(c, 1) * sizeof(T)), "hipMalloc");
        n = c;
    }
};

thread_local std::string g_nn_err;

}  // namespace

// Layer l: input channels cin(l) = 8 for l = 0, else 32; 32 output channels
// (the last layer's first 12 are used); dilation dil[l]; ReLU on all but the last.
struct m3d_net {
    hipStream_t stream = nullptr;
    m3d_context* ctx = nullptr;
    int max_batch = 0;
    int dil[NL] = {1, 2, 4, 1};
    int step = 0;
    Buf<float> w[NL], b[NL], gw[NL], gb[NL], mw[NL], vw[NL], mb[NL], vb[NL];
    Buf<unsigned short> wt[NL], wb[NL];
    Buf<unsigned short> x0, act[NL - 1], grad_a, grad_b, dout;
    Buf<float> out, part, lpart, gscale;
    float clip = 0.f;  // global gradient-norm clip (0: none)
    static int cin(int l) { return l == 0 ? CIN0 : HID; }
};

namespace {

void upload_bf16(m3d_net& n) {
    for (int l = 0; l < NL; ++l) {
        const int c = m3d_net::cin(l);
        hipLaunchKernelGGL(k_weights_bf16, dim3((27 * HID * c + 255) / 256), dim3(256), 0, n.stream, c,
                           n.w[l].p, n.wt[l].p, l > 0 ? n.wb[l].p : nullptr);
    }
    chk(hipGetLastError(), "k_weights_bf16");
}

void forward(m3d_net& n, int batch, const float* d_in, float* d_out) {
    const int nvox = batch * V;
    hipStream_t s = n.stream;
    hipLaunchKernelGGL(k_prep_input, dim3((nvox * CIN0 + 255) / 256), dim3(256), 0, s, nvox, d_in, n.x0.p);
    const dim3 grid((batch * G * G + 3) / 4), blk(256);
    hipLaunchKernelGGL((k_conv_fwd<CIN0, true, false>), grid, blk, 0, s, batch, n.dil[0], n.x0.p, n.wt[0].p,
                       n.b[0].p, n.act[0].p, nullptr, HID);
    // 32-channel layers: weights in LDS, two workgroups per CU.
    const dim3 gw(std::min((batch * G * G + kConvWaves - 1) / kConvWaves, 512)), bw(kConvWaves * 64);
    hipLaunchKernelGGL((k_conv_fwd_w<false, true, false>), gw, bw, kConvWBytes, s, batch, n.dil[1], n.act[0].p,
                       n.wt[1].p, n.b[1].p, n.act[1].p, nullptr, HID);
    hipLaunchKernelGGL((k_conv_fwd_w<false, true, false>), gw, bw, kConvWBytes, s, batch, n.dil[2], n.act[1].p,
                       n.wt[2].p, n.b[2].p, n.act[2].p, nullptr, HID);
    hipLaunchKernelGGL((k_conv_fwd_w<false, false, true>), gw, bw, kConvWBytes, s, batch, n.dil[3], n.act[2].p,
                       n.wt[3].p, n.b[3].p, nullptr, d_out, COUT);
    chk(hipGetLastError(), "conv forward");
}

// Loss partials (lpart) and parameter gradients (into gw / gb) for one
// batch, enqueued on the stream; read_loss waits for them.
constexpr int kLossBlocks = 512;
void backward(m3d_net& n, int batch, const float* d_in, const float* d_tgt) {
    const int nvox = batch * V;
    hipStream_t s = n.stream;
    forward(n, batch, d_in, n.out.p);
    hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, s, nvox, n.out.p, d_tgt,
                       1.0f / float(size_t(nvox) * COUT), n.dout.p, n.lpart.p);
    const dim3 grid((batch * G * G + 3) / 4), blk(256);
    const int rows = batch * G * G;
    const int nchunk = 64;
    const int rpc = (rows + nchunk - 1) / nchunk;
    // dz of layer l is grad (masked by act[l] for ReLU layers).
    const unsigned short* g = n.dout.p;
    unsigned short* bufs[2] = {n.grad_a.p, n.grad_b.p};
    const dim3 gw(std::min((batch * G * G + kConvWaves - 1) / kConvWaves, 512)), bw(kConvWaves * 64);
    for (int l = NL - 1; l >= 0; --l) {
        const unsigned short* a_in = (l == 0) ? n.x0.p : n.act[l - 1].p;
        // ReLU layers: dz = dA * 1[act > 0], formed once in place (dA is not
        // needed afterwards), so the kernels below read dz unmasked.
        if (l < NL - 1) {
            const size_t units = size_t(nvox) * HID / 8;
            hipLaunchKernelGGL(k_mask_relu, dim3(std::min<size_t>((units + 255) / 256, 4096)), dim3(256), 0, s, units,
                               const_cast<unsigned short*>(g), n.act[l].p);
        }
        const int c = m3d_net::cin(l);
        if (c == CIN0)
            hipLaunchKernelGGL((k_conv_bwd_w<CIN0>), dim3(nchunk, 9), blk, 0, s, batch, n.dil[l], rpc, a_in, g,
                               n.part.p);
        else
            hipLaunchKernelGGL((k_conv_bwd_w<HID>), dim3(nchunk, 9), blk, 0, s, batch, n.dil[l], rpc, a_in, g,
                               n.part.p);
        const int cnt = 27 * HID * c;
        hipLaunchKernelGGL(k_sum_chunks, dim3((cnt + 31) / 32), dim3(256), 0, s, nchunk, cnt, n.part.p,
                           n.gw[l].p);
        hipLaunchKernelGGL(k_bias_grad, dim3(kBiasBlocks), dim3(256), 0, s, nvox, g, nullptr, n.part.p);
        hipLaunchKernelGGL(k_sum_bias, dim3(1), dim3(1024), 0, s, kBiasBlocks, n.part.p, n.gb[l].p);
        if (l > 0) {
            unsigned short* dst = bufs[l & 1];
            hipLaunchKernelGGL((k_conv_fwd_w<true, false, false>), gw, bw, kConvWBytes, s, batch, n.dil[l], g,
                               n.wb[l].p, nullptr, dst, nullptr, HID);
            g = dst;
        }
    }
    chk(hipGetLastError(), "conv backward");
}

// The loss of the last backward(): waits for the stream.
float read_loss(m3d_net& n, int batch) {
    std::vector<float> lp(kLossBlocks);
    chk(hipMemcpyAsync(lp.data(), n.lpart.p, kLossBlocks * sizeof(float), hipMemcpyDeviceToHost, n.stream),
        "loss copy");
    chk(hipStreamSynchronize(n.stream), "sync");
    double tot = 0;
    for (float v : lp) tot += v;
    return float(tot / (double(batch) * V * COUT));
}

template <typename F>
int nn_guard(F&& f) {
    try {
        f();
        return M3D_OK;
    } catch (const std::exception& e) {
        g_nn_err = e.what();
        return M3D_ERR_HIP;
    }
}

}  // namespace

extern "C" {

const char* m3d_net_last_error(void) { return g_nn_err.c_str(); }

int m3d_net_create(m3d_context* ctx, int max_batch, const int* dilations, unsigned seed, m3d_net** out) {
    return nn_guard([&] {
        if (!ctx || !out || max_batch <= 0) throw std::invalid_argument("invalid arguments");
        std::unique_ptr<m3d_net> n(new m3d_net);
        n->ctx = ctx;
        n->stream = static_cast<hipStream_t>(m3d_context_stream(ctx));
        n->max_batch = max_batch;
        if (dilations) for (int l = 0; l < NL; ++l) n->dil[l] = std::max(1, dilations[l]);
        // He-uniform initialisation from a fixed LCG (deterministic per seed).
        unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
        auto urand = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return double((st >> 11) & ((1ull << 53) - 1)) / double(1ull << 53);
        };
        for (int l = 0; l < NL; ++l) {
            const int c = m3d_net::cin(l);
            const int cnt = 27 * HID * c;
            const int fan_in = 27 * (l == 0 ? CIN_USED : HID);
            const double lim = std::sqrt(6.0 / fan_in) * (l == NL - 1 ? 0.1 : 1.0);
            std::vector<float> w(cnt, 0.f), b(HID, 0.f);
            for (int i = 0; i < cnt; ++i) {
                const int co = (i / c) % HID, ci = i % c;
                const bool used = (l > 0 || ci < CIN_USED) && (l < NL - 1 || co < COUT);
                const double r = urand();
                w[i] = used ? float((2.0 * r - 1.0) * lim) : 0.f;
            }
            n->w[l].alloc(cnt); n->gw[l].alloc(cnt); n->mw[l].alloc(cnt); n->vw[l].alloc(cnt);
            n->b[l].alloc(HID); n->gb[l].alloc(HID); n->mb[l].alloc(HID); n->vb[l].alloc(HID);
            n->wt[l].alloc(cnt);
            if (l > 0) n->wb[l].alloc(cnt);
            chk(hipMemcpyAsync(n->w[l].p, w.data(), cnt * sizeof(float), hipMemcpyHostToDevice, n->stream), "w");
            chk(hipMemcpyAsync(n->b[l].p, b.data(), HID * sizeof(float), hipMemcpyHostToDevice, n->stream), "b");
            chk(hipMemsetAsync(n->mw[l].p, 0, cnt * sizeof(float), n->stream), "m");
            chk(hipMemsetAsync(n->vw[l].p, 0, cnt * sizeof(float), n->stream), "v");
            chk(hipMemsetAsync(n->mb[l].p, 0, HID * sizeof(float), n->stream), "m");
            chk(hipMemsetAsync(n->vb[l].p, 0, HID * sizeof(float), n->stream), "v");
        }
        const size_t nv = size_t(max_batch) * V;
        n->x0.alloc(nv * CIN0);
        for (int l = 0; l < NL - 1; ++l) n->act[l].alloc(nv * HID);
        n->grad_a.alloc(nv * HID);
        n->grad_b.alloc(nv * HID);
        n->dout.alloc(nv * HID);
        n->out.alloc(nv * COUT);
        n->part.alloc(size_t(64) * 27 * HID * HID + kBiasBlocks * 32);
        n->lpart.alloc(512);
        n->gscale.alloc(kSqBlocks);
        upload_bf16(*n);
        chk(hipStreamSynchronize(n->stream), "sync");
        *out = n.release();
    });
}

int m3d_net_destroy(m3d_net* n) {
    return nn_guard([&] {
        if (!n) return;
        (void)hipStreamSynchronize(n->stream);
        delete n;
    });
}

// Parameter count: weights 27*32*(8 + 3*32) floats in layer order, each
// [tap][co][ci]; biases 4*32.
int m3d_net_num_params(long long* nw, long long* nb) {
    long long w = 0;
    for (int l = 0; l < NL; ++l) w += 27LL * HID * m3d_net::cin(l);
    if (nw) *nw = w;
    if (nb) *nb = NL * HID;
    return M3D_OK;
}

int m3d_net_get_params(m3d_net* n, float* w, float* b, float* gw, float* gb) {
    return nn_guard([&] {
        size_t ow = 0;
        for (int l = 0; l < NL; ++l) {
            const size_t cnt = size_t(27) * HID * m3d_net::cin(l);
            if (w) chk(hipMemcpyAsync(w + ow, n->w[l].p, cnt * 4, hipMemcpyDeviceToHost, n->stream), "get w");
            if (gw) chk(hipMemcpyAsync(gw + ow, n->gw[l].p, cnt * 4, hipMemcpyDeviceToHost, n->stream), "get gw");
            if (b) chk(hipMemcpyAsync(b + l * HID, n->b[l].p, HID * 4, hipMemcpyDeviceToHost, n->stream), "get b");
            if (gb) chk(hipMemcpyAsync(gb + l * HID, n->gb[l].p, HID * 4, hipMemcpyDeviceToHost, n->stream), "get gb");
            ow += cnt;
        }
        chk(hipStreamSynchronize(n->stream), "sync");
    });
}

int m3d_net_set_params(m3d_net* n, const float* w, const float* b) {
    return nn_guard([&] {
        size_t ow = 0;
        for (int l = 0; l < NL; ++l) {
            const size_t cnt = size_t(27) * HID * m3d_net::cin(l);
            chk(hipMemcpyAsync(n->w[l].p, w + ow, cnt * 4, hipMemcpyHostToDevice, n->stream), "set w");
            chk(hipMemcpyAsync(n->b[l].p, b + l * HID, HID * 4, hipMemcpyHostToDevice, n->stream), "set b");
            ow += cnt;
        }
        upload_bf16(*n);
        chk(hipStreamSynchronize(n->stream), "sync");
    });
}

// d_in: device [batch][32^3][2] FP32; d_out: device [batch][32^3][12] FP32.
int m3d_net_forward(m3d_net* n, int batch, const float* d_in, float* d_out) {
    return nn_guard([&] {
        if (batch <= 0 || batch > n->max_batch) throw std::invalid_argument("batch out of range");
        forward(*n, batch, d_in, d_out);
        chk(hipStreamSynchronize(n->stream), "sync");
    });
}

// Loss and gradients (readable with m3d_net_get_params) without an update.
int m3d_net_gradients(m3d_net* n, int batch, const float* d_in, const float* d_tgt, float* loss) {
    return nn_guard([&] {
        if (batch <= 0 || batch > n->max_batch) throw std::invalid_argument("batch out of range");
        backward(*n, batch, d_in, d_tgt);
        const float l = read_loss(*n, batch);
        if (loss) *loss = l;
    });
}

int m3d_net_prolong(m3d_net* n, int batch, const float* d_field, int n_edges, const float* d_edge_xyz,
                    const unsigned char* d_edge_dir, const double* box, double* d_x0) {
    return nn_guard([&] {
        if (batch <= 0 || n_edges <= 0 || !box || !d_field || !d_edge_xyz || !d_edge_dir || !d_x0)
            throw std::invalid_argument("invalid arguments");
        const float hx = float((box[1] - box[0]) / G), hy = float((box[3] - box[2]) / G),
                    hz = float((box[5] - box[4]) / G);
        const size_t total = size_t(batch) * n_edges;
        const int grid = int(std::min<size_t>((total + 255) / 256, 8192));
        hipLaunchKernelGGL(k_prolong, dim3(grid), dim3(256), 0, n->stream, batch, n_edges, d_field, d_edge_xyz,
                           d_edge_dir, float(box[0]), float(box[2]), float(box[4]), hx, hy, hz,
                           reinterpret_cast<double2*>(d_x0));
        chk(hipGetLastError(), "k_prolong");
        chk(hipStreamSynchronize(n->stream), "sync");
    });
}

int m3d_net_set_clip(m3d_net* n, float max_norm) {
    return nn_guard([&] {
        if (!n || !(max_norm >= 0.f)) throw std::invalid_argument("invalid arguments");
        n->clip = max_norm;
    });
}

int m3d_net_train_step(m3d_net* n, int batch, const float* d_in, const float* d_tgt, float lr, float* loss) {
    return nn_guard([&] {
        if (batch <= 0 || batch > n->max_batch) throw std::invalid_argument("batch out of range");
        backward(*n, batch, d_in, d_tgt);
        n->step++;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f - std::pow(b1, float(n->step)), c2 = 1.f - std::pow(b2, float(n->step));
        ParamSet P;
        int off = 0;
        for (int k = 0; k < NL; ++k) {
            const int c = m3d_net::cin(k);
            P.p[k] = n->w[k].p; P.g[k] = n->gw[k].p; P.m[k] = n->mw[k].p; P.v[k] = n->vw[k].p;
            P.p[NL + k] = n->b[k].p; P.g[NL + k] = n->gb[k].p; P.m[NL + k] = n->mb[k].p; P.v[NL + k] = n->vb[k].p;
            P.wt[k] = n->wt[k].p;
            P.wb[k] = k > 0 ? n->wb[k].p : nullptr;
            P.cin[k] = c;
            P.off[k] = off;
            off += 27 * HID * c;
        }
        for (int k = 0; k < NL; ++k) {
            P.off[NL + k] = off;
            off += HID;
        }
        P.off[2 * NL] = off;
        hipLaunchKernelGGL(k_grad_sq, dim3(kSqBlocks), dim3(256), 0, n->stream, P, n->gscale.p);
        hipLaunchKernelGGL(k_adam_all, dim3((off + 255) / 256), dim3(256), 0, n->stream, P, lr, b1, b2, eps, c1, c2,
                           n->clip, n->gscale.p);
        chk(hipGetLastError(), "adam");
        const float l = read_loss(*n, batch);
        if (loss) *loss = l;
    });
}

}  // extern "C"
