// Host orchestration of the device solve and the C-ABI declared in
// include/m3d_device.h.
//
// An operator is laid out once per mesh: rows are greedily coloured on the
// free-free coupling graph, renumbered colour by colour (each colour padded
// to a 64-row slice), and every slice's lower / upper couplings are packed
// as SELL-64 blocks.  In that order the natural-order symmetric Gauss-Seidel
// of ssor.h becomes a multicolour sweep: one launch per colour, every row of
// a colour independent.  BiCGStab scalars live on the device; the host only
// polls the per-system status between chunks of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "abi_common.h"
#include "kernels.h"
#include "m3d_device.h"

using namespace m3d::dev;

namespace m3d {
namespace abi {

std::string& last_error() {
    thread_local std::string msg;
    return msg;
}

std::mutex& context_mutex() {
    static std::mutex m;
    return m;
}

void context_free_if_unused(m3d_context* ctx) {  // with the lock held
    if (!ctx->closed || ctx->operators > 0) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void context_attach(m3d_context* ctx) {
    std::lock_guard<std::mutex> g(context_mutex());
    ctx->operators++;
}

void context_detach(m3d_context* ctx) {
    std::lock_guard<std::mutex> g(context_mutex());
    ctx->operators--;
    context_free_if_unused(ctx);
}

void context_close(m3d_context* ctx) {
    std::lock_guard<std::mutex> g(context_mutex());
    ctx->closed = true;
    context_free_if_unused(ctx);
}

}  // namespace abi
}  // namespace m3d

using namespace m3d::abi;

// Device scratch of the allocation-free sensitivity entry point
// (m3d_sensitivity_device_noalloc): the frequencies of one call, staged on
// the device.  Each GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: it uses the address hipGetSymbolAddress returns on each device.
constexpr int kNoAllocMaxSystems = 4096;
static __device__ __align__(256) double scratch[kNoAllocMaxSystems];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Address of this device's scratch; looked up once per device, with the
// device current (the caller has set it).
double* device_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Fail(M3D_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Fail(M3D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    return g_scratch_ptr[device];
}

}  // namespace

namespace {

// Per-(batch width, precision) device workspace of one operator.
template <int NS, typename V>
struct Work {
    using C = cvec<V>;
    DBuf<C> b, x, r, r0, p, v, s, t, y, ph, sh;
};

// Instantiated HIP graphs of fixed-length iteration chunks, keyed by the
// solution vector they update and the chunk kind.  Kernel arguments only
// reference buffers that live as long as the workspace.
struct GraphCache {
    struct Entry {
        const void* key;
        int kind;
        int len;
        hipGraphExec_t exec;
    };
    std::vector<Entry> entries;
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() { clear(); }
    void clear() {
        for (auto& e : entries) (void)hipGraphExecDestroy(e.exec);
        entries.clear();
    }
    template <typename F>
    hipGraphExec_t get(hipStream_t s, const void* key, int kind, int len, F&& body) {
        for (auto& e : entries)
            if (e.key == key && e.kind == kind && e.len == len) return e.exec;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) throw Fail(M3D_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        entries.push_back({key, kind, len, ex});
        return ex;
    }
};

bool use_graphs() {
    static const bool on = [] {
        const char* v = std::getenv("M3D_GRAPHS");
        return !(v && v[0] == '0');
    }();
    return on;
}

// A p_hat and A s_hat by Eisenstat's identity (lower couplings + D y) unless
// M3D_EISENSTAT=0 asks for the full product.
bool use_eisenstat() {
    static const bool on = [] {
        const char* v = std::getenv("M3D_EISENSTAT");
        return !(v && v[0] == '0');
    }();
    return on;
}

// Largest value table of packed SELL blocks (32 KB of doubles).
constexpr int kMaxValueTable = 4096;

uint64_t bits_of(double a) {
    uint64_t b;
    std::memcpy(&b, &a, sizeof b);
    return b;
}

// Bit pattern of a double -> its index in the value table: open addressing
// over twice kMaxValueTable slots (the table never holds more keys).
class FlatIndex {
  public:
    FlatIndex() : key_(kSlots, 0), val_(kSlots, -1) {}
    // Adds key -> v unless present; true when added.
    bool insert(uint64_t key, int v) {
        size_t h = slot(key);
        while (val_[h] >= 0) {
            if (key_[h] == key) return false;
            h = (h + 1) & (kSlots - 1);
        }
        key_[h] = key;
        val_[h] = v;
        return true;
    }
    int find(uint64_t key) const {
        for (size_t h = slot(key); val_[h] >= 0; h = (h + 1) & (kSlots - 1))
            if (key_[h] == key) return val_[h];
        return -1;
    }

  private:
    static constexpr size_t kSlots = 4 * kMaxValueTable;
    static size_t slot(uint64_t k) {
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdULL;
        k ^= k >> 33;
        return size_t(k) & (kSlots - 1);
    }
    std::vector<uint64_t> key_;
    std::vector<int> val_;
};

// The last colour's backward step folded into its forward launch unless
// M3D_FUSE_LAST_COLOUR=0 (read at each operator build).
bool fuse_last_colour() {
    const char* v = std::getenv("M3D_FUSE_LAST_COLOUR");
    return !(v && v[0] == '0');
}

// Packed SELL blocks unless M3D_PACKED_SELL=0 (read at each operator build).
bool packed_sell() {
    const char* v = std::getenv("M3D_PACKED_SELL");
    return !(v && v[0] == '0');
}

// Row tiles for the LDS-staged plain product at widths 8-32 unless
// M3D_TILED_SPMV=0 (read at each operator build).
bool tiled_spmv() {
    const char* v = std::getenv("M3D_TILED_SPMV");
    return !(v && v[0] == '0');
}

// LDS a tile may stage per workgroup: 63 KB, or M3D_TILE_STAGE_BYTES
// (8192-64512), and never more than the 64 KB a workgroup's launch may ask
// for beside the value table of ntab doubles (16-byte aligned, as the
// kernels lay it out).
int tile_stage_bytes(int ntab) {
    const char* v = std::getenv("M3D_TILE_STAGE_BYTES");
    const int b = std::min(63 * 1024, std::max(8192, v ? std::atoi(v) : 63 * 1024));
    const int table = (ntab * int(sizeof(double)) + 15) / 16 * 16;
    return std::min(b, 64 * 1024 - table);
}

// Entries of every device row in the order of its SELL blocks (lower
// couplings, then upper): columns as device rows, values as table indices.
struct RowEntries {
    std::vector<int> ptr, col, tab;
};

// Spatial row-block order for the plain SpMV (M3D_SPMV_ORDER=0: device
// order).  The order changes which rows run together, never a row's values.
bool spmv_order() {
    const char* v = std::getenv("M3D_SPMV_ORDER");
    return !(v && v[0] == '0');
}

bool row_sort() {
    static const bool on = [] {
        const char* v = std::getenv("M3D_ROW_SORT");
        return v && v[0] == '1';
    }();
    return on;
}

bool use_compaction() {
    static const bool on = [] {
        const char* v = std::getenv("M3D_COMPACT");
        return !(v && v[0] == '0');
    }();
    return on;
}

constexpr int kGraphIters = 8;     // BiCGStab iterations per graph launch
constexpr int kGraphPcgIters = 16; // DivCorr PCG iterations per graph launch

template <int NS>
struct Work64 {
    GraphCache graphs;
    DBuf<double2> b, x, aux, tmp, rres;             // FP64 system vectors
    DBuf<double2> phi0, dr, dz, dp, dphi, dap;      // DivCorr node vectors
    DBuf<double> omega, scale;
    std::vector<double> h_omega;  // host copy of omega
    DBuf<SysState> st;
    DBuf<double> part;
    DBuf<int> flags;
    std::unique_ptr<Work<NS, double>> w64;
    std::unique_ptr<Work<NS, float>> w32;
    bool ready = false;
};

}  // namespace

struct m3d_operator {
    m3d_context* ctx = nullptr;
    int n = 0;          // caller edges
    int npad = 0;       // device rows
    int ncolours = 0;
    long long stored = 0, nnz = 0;
    std::vector<int> perm, iperm, colour_start;
    std::vector<int> hang_level_start;  // offsets into the hanging triples per level
    // Operator arrays.
    DBuf<int> d_perm, l_col, u_col, hang;
    DBuf<int4> slc;
    DBuf<double> l_val, u_val, diag_s, me;
    DBuf<float> l_val32, u_val32, diag_s32, me32;
    // Packed SELL blocks and their value table (empty when the table does not
    // fit beside the column bits, or packing is switched off).
    DBuf<unsigned> l_pk, u_pk;
    DBuf<double> vtab;
    DBuf<float> vtab32;
    int cbits = 0, ntab = 0;
    // The last colour has no upper couplings (its backward step is z = y).
    bool last_upper_empty = false;
    // Row blocks per batch width in spatial order (empty: device order).
    DBuf<int> order[kWidths];
    // Row tiles of the LDS-staged plain product per batch width.
    struct Tiles {
        int ntiles = 0, lbits = 0, max_stage = 0;
        DBuf<int> stage_off, stage_rows, slice_off;
        DBuf<int4> slc;
        DBuf<int2> slot;
        DBuf<unsigned> words;
    } tiles[kWidths];
    DBuf<unsigned char> bc, constrained;
    std::vector<unsigned char> h_constrained;
    bool mass_set = false;
    // DivCorr arrays.
    bool dc_set = false;
    int nn = 0;
    DBuf<int> nt_ptr, nt_edge, g_col, l_ptr, l_colidx;
    DBuf<double> nt_val, g_val, ms, l_vals, l_diag;
    DBuf<unsigned char> dc_bc_edge, dc_dead, dc_bc_node;
    // FP32 copies of the Laplacian for the FP32 PCG, made on first use after
    // the values change (l32_fresh false).
    DBuf<float> l_vals32, l_diag32;
    bool l32_fresh = false;
    // Sigma update: edge-side mass structure (caller edge order), the
    // Laplacian entries' edges (device rows) and which edges its pattern holds.
    bool mass_struct = false, sigma_update_ok = false;
    int mass_nc = 0;
    DBuf<int> m_adj_ptr, m_adj_cell, l_edge, d_err;
    DBuf<double> m_area, m_len;
    DBuf<unsigned char> m_massless, m_dead, dc_in_pattern;
    // Sensitivity arrays.
    int nc = 0;
    DBuf<int> s_cell_ptr, s_cell_edges;
    DBuf<int> s_edge_ptr, s_edge_cells;   // edge -> cells containing it, ascending
    DBuf<double> s_weight, s_sigma, s_omega;
    // Launch shapes by precision (0 = complex128, 1 = complex64) and width.
    LaunchShape shape[2][kWidths];
    // Workspaces by batch width.
    std::unique_ptr<Work64<1>> w1;
    std::unique_ptr<Work64<2>> w2;
    std::unique_ptr<Work64<4>> w4;
    std::unique_ptr<Work64<8>> w8;
    std::unique_ptr<Work64<16>> w16;
    std::unique_ptr<Work64<32>> w32;

    template <typename V>
    DevOp<V> view() const;
    void set_tiles(DevTiles* t) const {
        for (int i = 0; i < kWidths; ++i) {
            const Tiles& h = tiles[i];
            t[i].ntiles = h.ntiles;
            t[i].lbits = h.lbits;
            t[i].max_stage = h.max_stage;
            t[i].stage_off = h.stage_off.p;
            t[i].stage_rows = h.stage_rows.p;
            t[i].slice_off = h.slice_off.p;
            t[i].slc = h.slc.p;
            t[i].slot = h.slot.p;
            t[i].words = h.words.p;
        }
    }
};

template <>
DevOp<double> m3d_operator::view<double>() const {
    DevOp<double> o;
    o.n = npad;
    for (int i = 0; i < kWidths; ++i) o.shape[i] = shape[0][i];
    o.slc = slc.p;
    o.l_val = l_val.p; o.l_col = l_col.p; o.u_val = u_val.p; o.u_col = u_col.p;
    o.l_pk = l_pk.p; o.u_pk = u_pk.p; o.vtab = ntab ? vtab.p : nullptr; o.cbits = cbits; o.ntab = ntab;
    o.diag_s = diag_s.p; o.me = me.p;
    for (int i = 0; i < kWidths; ++i) o.order[i] = order[i].p;
    set_tiles(o.tiles);
    return o;
}
template <>
DevOp<float> m3d_operator::view<float>() const {
    DevOp<float> o;
    o.n = npad;
    for (int i = 0; i < kWidths; ++i) o.shape[i] = shape[1][i];
    o.slc = slc.p;
    o.l_val = l_val32.p; o.l_col = l_col.p; o.u_val = u_val32.p; o.u_col = u_col.p;
    o.l_pk = l_pk.p; o.u_pk = u_pk.p; o.vtab = ntab ? vtab32.p : nullptr; o.cbits = cbits; o.ntab = ntab;
    o.diag_s = diag_s32.p; o.me = me32.p;
    for (int i = 0; i < kWidths; ++i) o.order[i] = order[i].p;
    set_tiles(o.tiles);
    return o;
}

namespace {

template <int NS>
Work64<NS>& workspace(m3d_operator& op);
template <> Work64<1>& workspace<1>(m3d_operator& op) { if (!op.w1) op.w1.reset(new Work64<1>); return *op.w1; }
template <> Work64<2>& workspace<2>(m3d_operator& op) { if (!op.w2) op.w2.reset(new Work64<2>); return *op.w2; }
template <> Work64<4>& workspace<4>(m3d_operator& op) { if (!op.w4) op.w4.reset(new Work64<4>); return *op.w4; }
template <> Work64<8>& workspace<8>(m3d_operator& op) { if (!op.w8) op.w8.reset(new Work64<8>); return *op.w8; }
template <> Work64<16>& workspace<16>(m3d_operator& op) { if (!op.w16) op.w16.reset(new Work64<16>); return *op.w16; }
template <> Work64<32>& workspace<32>(m3d_operator& op) { if (!op.w32) op.w32.reset(new Work64<32>); return *op.w32; }

template <int NS>
Work64<NS>& ready_workspace(m3d_operator& op) {
    Work64<NS>& W = workspace<NS>(op);
    if (W.ready) return W;
    const size_t nv = size_t(op.npad) * NS;
    W.b.alloc(nv); W.x.alloc(nv); W.aux.alloc(nv); W.tmp.alloc(nv); W.rres.alloc(nv);
    W.omega.alloc(NS); W.scale.alloc(NS);
    W.st.alloc(kMaxSys);
    W.part.alloc(size_t(4) * kMaxSys * kMaxBlocks);
    W.flags.alloc(NS);
    HIPCHK(hipMemsetAsync(W.st.p, 0, kMaxSys * sizeof(SysState), op.ctx->stream));
    W.ready = true;
    return W;
}

template <int NS, typename V>
Work<NS, V>& vec_workspace(m3d_operator& op, Work64<NS>& W);
template <int NS>
Work<NS, double>& vec64(m3d_operator& op, Work64<NS>& W) {
    if (!W.w64) {
        W.w64.reset(new Work<NS, double>);
        const size_t nv = size_t(op.npad) * NS;
        auto& w = *W.w64;
        w.r.alloc(nv); w.r0.alloc(nv); w.p.alloc(nv); w.v.alloc(nv); w.s.alloc(nv);
        w.t.alloc(nv); w.y.alloc(nv); w.ph.alloc(nv); w.sh.alloc(nv);
        hipStream_t st = op.ctx->stream;
        for (auto* b : {&w.r, &w.r0, &w.p, &w.v, &w.s, &w.t, &w.y, &w.ph, &w.sh})
            Launch<NS, double>::fill_zero(st, nv, b->p);
    }
    return *W.w64;
}
template <int NS>
Work<NS, float>& vec32(m3d_operator& op, Work64<NS>& W) {
    if (!W.w32) {
        W.w32.reset(new Work<NS, float>);
        const size_t nv = size_t(op.npad) * NS;
        auto& w = *W.w32;
        w.b.alloc(nv); w.x.alloc(nv);
        w.r.alloc(nv); w.r0.alloc(nv); w.p.alloc(nv); w.v.alloc(nv); w.s.alloc(nv);
        w.t.alloc(nv); w.y.alloc(nv); w.ph.alloc(nv); w.sh.alloc(nv);
        hipStream_t st = op.ctx->stream;
        for (auto* b : {&w.b, &w.x, &w.r, &w.r0, &w.p, &w.v, &w.s, &w.t, &w.y, &w.ph, &w.sh})
            Launch<NS, float>::fill_zero(st, nv, b->p);
    }
    return *W.w32;
}

template <int NS>
void ensure_dc_vectors(m3d_operator& op, Work64<NS>& W) {
    const size_t nv = size_t(std::max(op.nn, 1)) * NS;
    if (W.dr.n >= nv) return;
    W.phi0.alloc(nv); W.dr.alloc(nv); W.dz.alloc(nv); W.dp.alloc(nv); W.dphi.alloc(nv); W.dap.alloc(nv);
}

std::vector<SysState> read_states(m3d_operator& op, const DBuf<SysState>& st) {
    std::vector<SysState> h(kMaxSys);
    HIPCHK(hipMemcpyAsync(h.data(), st.p, kMaxSys * sizeof(SysState), hipMemcpyDeviceToHost,
                          op.ctx->stream));
    HIPCHK(hipStreamSynchronize(op.ctx->stream));
    return h;
}

// One BiCGStab iteration (bicgstab.cpp:50-160) for every running system.
// One multicolour SSOR application z = (D+U)^-1 D (D+L)^-1 in: forward
// sweeps over the colours in order (mode: the input as given, or formed from
// in = r and written to p by the search-direction or s update), then
// backward sweeps in reverse order.  The last colour couples only to earlier colours, so when its
// upper block is empty its backward step z = y is written by its forward
// launch and the backward sweeps start one colour earlier.
template <int NS, typename V>
void ssor_sweeps(hipStream_t s, const m3d_operator& op, const DevOp<V>& A, const double* omega, int mode,
                 const cvec<V>* in, cvec<V>* p, const cvec<V>* v, cvec<V>* y, cvec<V>* z,
                 const SysState* st) {
    using L = Launch<NS, V>;
    const int K = op.ncolours;
    const bool fuse = op.last_upper_empty;
    for (int c = 0; c < K; ++c)
        L::sweep_fwd(s, A, omega, op.colour_start[c], op.colour_start[c + 1], mode, in, p, v, y, st,
                     fuse && c == K - 1 ? z : nullptr);
    for (int c = K - (fuse ? 2 : 1); c >= 0; --c)
        L::sweep_bwd(s, A, omega, op.colour_start[c], op.colour_start[c + 1], y, z, st);
}

template <int NS, typename V>
void bicg_iteration(m3d_operator& op, Work64<NS>& W, Work<NS, V>& w, cvec<V>* x) {
    hipStream_t s = op.ctx->stream;
    const DevOp<V> A = op.view<V>();
    using L = Launch<NS, V>;
    ssor_sweeps<NS, V>(s, op, A, W.omega.p, kInPUpdate, w.r.p, w.p.p, w.v.p, w.y.p, w.ph.p, W.st.p);
    // v = A p_hat; in Eisenstat form from y, the forward-sweep output of p.
    const bool eis = use_eisenstat();
    L::spmv_v(s, A, W.omega.p, w.ph.p, eis ? w.y.p : nullptr, w.r0.p, w.v.p, W.st.p, W.part.p);
    // s = r - alpha v is formed by the forward sweeps of s_hat = M^-1 s; its
    // norm test follows in the (t, t), (t, s) reduction.
    ssor_sweeps<NS, V>(s, op, A, W.omega.p, kInSUpdate, w.r.p, w.s.p, w.v.p, w.y.p, w.sh.p, W.st.p);
    L::spmv_t(s, A, W.omega.p, w.sh.p, eis ? w.y.p : nullptr, w.s.p, w.t.p, W.st.p, W.part.p);
    L::bicg_xr(s, op.npad, x, w.ph.p, w.sh.p, w.s.p, w.t.p, w.r0.p, w.r.p, W.st.p, W.part.p);
}

struct BicgOutcome {
    int iterations[kMaxSys];
    double residual[kMaxSys];
    int status[kMaxSys];
    int info[kMaxSys];
};

template <int NS, typename V>
Work<NS, V>& vecws(m3d_operator& op, Work64<NS>& W);
template <> Work<1, double>& vecws<1, double>(m3d_operator& op, Work64<1>& W) { return vec64<1>(op, W); }
template <> Work<2, double>& vecws<2, double>(m3d_operator& op, Work64<2>& W) { return vec64<2>(op, W); }
template <> Work<4, double>& vecws<4, double>(m3d_operator& op, Work64<4>& W) { return vec64<4>(op, W); }
template <> Work<8, double>& vecws<8, double>(m3d_operator& op, Work64<8>& W) { return vec64<8>(op, W); }
template <> Work<16, double>& vecws<16, double>(m3d_operator& op, Work64<16>& W) { return vec64<16>(op, W); }
template <> Work<32, double>& vecws<32, double>(m3d_operator& op, Work64<32>& W) { return vec64<32>(op, W); }
template <> Work<1, float>& vecws<1, float>(m3d_operator& op, Work64<1>& W) { return vec32<1>(op, W); }
template <> Work<2, float>& vecws<2, float>(m3d_operator& op, Work64<2>& W) { return vec32<2>(op, W); }
template <> Work<4, float>& vecws<4, float>(m3d_operator& op, Work64<4>& W) { return vec32<4>(op, W); }
template <> Work<8, float>& vecws<8, float>(m3d_operator& op, Work64<8>& W) { return vec32<8>(op, W); }
template <> Work<16, float>& vecws<16, float>(m3d_operator& op, Work64<16>& W) { return vec32<16>(op, W); }
template <> Work<32, float>& vecws<32, float>(m3d_operator& op, Work64<32>& W) { return vec32<32>(op, W); }

// Solution buffer of a narrow workspace for continued solves.
template <int NB, typename V>
cvec<V>* narrow_x(m3d_operator& op, Work64<NB>& W);
template <int NB, typename V>
struct NarrowX;
template <int NB>
struct NarrowX<NB, double> {
    static double2* get(m3d_operator&, Work64<NB>& W) { return W.rres.p; }
};
template <int NB>
struct NarrowX<NB, float> {
    static float2* get(m3d_operator& op, Work64<NB>& W) { return vec32<NB>(op, W).x.p; }
};

template <int NS, typename V>
void bicg_loop(m3d_operator& op, Work64<NS>& W, Work<NS, V>& w, cvec<V>* x,
               std::vector<SysState>& h, int maxiter);

// Continue the running systems listed in map (wide system index per narrow
// slot) at batch width NB: move the iteration vectors (r, r0, p, v, x), the
// frequencies and the device state, iterate, and move x and the state back.
template <int NB, typename V>
void continue_narrow(m3d_operator& op, int NA, const std::vector<int>& map, Work<kMaxSys, V>* /*unused*/,
                     cvec<V>* r, cvec<V>* r0, cvec<V>* p, cvec<V>* v, cvec<V>* x,
                     const std::vector<double>& omega_wide, std::vector<SysState>& h_wide, int maxiter) {
    Work64<NB>& WB = ready_workspace<NB>(op);
    Work<NB, V>& wb = vecws<NB, V>(op, WB);
    cvec<V>* xb = NarrowX<NB, V>::get(op, WB);
    hipStream_t s = op.ctx->stream;
    std::vector<int> m(NB, -1);
    std::vector<double> om(NB, 0.0);
    std::vector<SysState> hs(kMaxSys);
    for (int j = 0; j < kMaxSys; ++j) hs[j] = SysState{}, hs[j].status = kIdle;
    for (int j = 0; j < int(map.size()); ++j) {
        m[j] = map[j];
        om[j] = omega_wide[map[j]];
        hs[j] = h_wide[map[j]];
    }
    WB.h_omega = om;
    HIPCHK(hipMemcpyAsync(WB.omega.p, om.data(), NB * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(WB.flags.p, m.data(), NB * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(WB.st.p, hs.data(), kMaxSys * sizeof(SysState), hipMemcpyHostToDevice, s));
    const int rows = op.npad;
    launch_regroup<V>(s, rows, NA, NB, WB.flags.p, r, wb.r.p, true);
    launch_regroup<V>(s, rows, NA, NB, WB.flags.p, r0, wb.r0.p, true);
    launch_regroup<V>(s, rows, NA, NB, WB.flags.p, p, wb.p.p, true);
    launch_regroup<V>(s, rows, NA, NB, WB.flags.p, v, wb.v.p, true);
    launch_regroup<V>(s, rows, NA, NB, WB.flags.p, x, xb, true);
    HIPCHK(hipGetLastError());
    bicg_loop<NB, V>(op, WB, wb, xb, hs, maxiter);
    launch_regroup<V>(s, rows, NA, NB, WB.flags.p, x, xb, false);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    for (int j = 0; j < int(map.size()); ++j) h_wide[map[j]] = hs[j];
}

// Iterate until every system in h has stopped; when few systems are still
// running, continue them at a narrower batch width.
template <int NS, typename V>
void bicg_loop(m3d_operator& op, Work64<NS>& W, Work<NS, V>& w, cvec<V>* x,
               std::vector<SysState>& h, int maxiter) {
    hipStream_t s = op.ctx->stream;
    int launched = 0;
    int chunk = 4;
    auto running = [&]() {
        int k = 0;
        for (int i = 0; i < NS; ++i)
            if (h[i].status == kRunning) ++k;
        return k;
    };
    hipGraphExec_t ex = nullptr;
    if (use_graphs())
        ex = W.graphs.get(s, x, sizeof(V), kGraphIters, [&] {
            for (int k = 0; k < kGraphIters; ++k) bicg_iteration<NS, V>(op, W, w, x);
        });
    // Iterations past a system's stop are no-ops for it (kernels check its
    // status on the device), so whole chunks can be replayed.
    while (launched < maxiter) {
        const int k = running();
        if (k == 0) break;
        if (NS > 1 && use_compaction() && k <= NS / 2) {
            std::vector<int> map;
            for (int i = 0; i < NS; ++i)
                if (h[i].status == kRunning) map.push_back(i);
            const int left = maxiter;  // per-system maxiter is enforced on the device
            if (k == 1) continue_narrow<1, V>(op, NS, map, nullptr, w.r.p, w.r0.p, w.p.p, w.v.p, x, W.h_omega, h, left);
            else if (k == 2) continue_narrow<2, V>(op, NS, map, nullptr, w.r.p, w.r0.p, w.p.p, w.v.p, x, W.h_omega, h, left);
            else if (k <= 4) continue_narrow<4, V>(op, NS, map, nullptr, w.r.p, w.r0.p, w.p.p, w.v.p, x, W.h_omega, h, left);
            else if (k <= 8) continue_narrow<8, V>(op, NS, map, nullptr, w.r.p, w.r0.p, w.p.p, w.v.p, x, W.h_omega, h, left);
            else continue_narrow<16, V>(op, NS, map, nullptr, w.r.p, w.r0.p, w.p.p, w.v.p, x, W.h_omega, h, left);
            HIPCHK(hipMemcpyAsync(W.st.p, h.data(), kMaxSys * sizeof(SysState), hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
            return;
        }
        if (ex) {
            HIPCHK(hipGraphLaunch(ex, s));
            launched += kGraphIters;
        } else {
            const int todo = std::min(chunk, maxiter - launched);
            for (int i = 0; i < todo; ++i) bicg_iteration<NS, V>(op, W, w, x);
            HIPCHK(hipGetLastError());
            launched += todo;
            chunk = std::min(chunk * 2, 32);
        }
        h = read_states(op, W.st);
    }
}

// Runs BiCGStab on the masked systems until every one has stopped.
template <int NS, typename V>
BicgOutcome bicgstab(m3d_operator& op, Work64<NS>& W, Work<NS, V>& w, const cvec<V>* b,
                     cvec<V>* x, unsigned mask, int maxiter, double tol, bool x_zero) {
    hipStream_t s = op.ctx->stream;
    launch_bicg_begin(s, W.st.p, mask, maxiter, tol);
    Launch<NS, V>::bicg_init(s, op.view<V>(), W.omega.p, b, x, w.r.p, w.r0.p, W.st.p, W.part.p,
                             x_zero ? 1 : 0);
    HIPCHK(hipGetLastError());
    std::vector<SysState> h = read_states(op, W.st);
    bicg_loop<NS, V>(op, W, w, x, h, maxiter);
    BicgOutcome o{};
    for (int i = 0; i < NS; ++i) {
        o.iterations[i] = h[i].iter;
        o.residual[i] = h[i].residual;
        o.status[i] = h[i].status;
        o.info[i] = h[i].info;
    }
    return o;
}

template <int NS>
void divcorr(m3d_operator& op, Work64<NS>& W, double2* E, const double2* phi0, unsigned mask,
             const m3d_solve_params& prm);

// The divergence-correction potential solve stays FP64.  On request
// (M3D_DIVCORR_FP32=1, read at each call) the DivCorr steps between the
// bursts of the mixed-precision path run their PCG in FP32 (complex64 node
// vectors, FP32 Laplacian, double reductions) when its tolerance is at least
// 1e-5, where FP32 rounding (6e-8) stays far below the stop.  The FP64 paths,
// the post-hoc correction and m3d_divcorr (1e-10) always run in FP64.
constexpr double kDcFp32MinTol = 1e-5;

bool dc_in_fp32(const m3d_solve_params& prm) {
    if (prm.precision != M3D_PREC_MIXED || !(prm.divcorr_pcg_tol >= kDcFp32MinTol)) return false;
    const char* e = getenv("M3D_DIVCORR_FP32");
    return e && e[0] == '1';
}

void ensure_laplacian32(m3d_operator& op) {
    if (op.l32_fresh) return;
    hipStream_t s = op.ctx->stream;
    // Reallocated only when the sizes change, which only set_divcorr does
    // (it clears the captured graphs that hold the old addresses).
    if (op.l_vals32.n != op.l_vals.n) op.l_vals32.alloc(op.l_vals.n);
    if (op.l_diag32.n != op.l_diag.n) op.l_diag32.alloc(op.l_diag.n);
    launch_to_f32(s, op.l_vals.n, op.l_vals.p, op.l_vals32.p);
    launch_to_f32(s, op.l_diag.n, op.l_diag.p, op.l_diag32.p);
    HIPCHK(hipGetLastError());
    op.l32_fresh = true;
}

// The DivCorr loop of divcorr_poisson_ (forward_solver_fv.cpp:623-722) with
// its PCG in V; the node vectors of W hold cvec<V> (float2 in the first half
// of each complex128 buffer for V = float).
template <int NS, typename V>
void divcorr_pcg(m3d_operator& op, Work64<NS>& W, double2* E, const double2* phi0, unsigned mask,
                 const m3d_solve_params& prm, const V* lval, const V* ldiag) {
    hipStream_t s = op.ctx->stream;
    using D = LaunchDc<NS, V>;
    using C = cvec<V>;
    C* dr = reinterpret_cast<C*>(W.dr.p);
    C* dz = reinterpret_cast<C*>(W.dz.p);
    C* dp = reinterpret_cast<C*>(W.dp.p);
    C* dphi = reinterpret_cast<C*>(W.dphi.p);
    C* dap = reinterpret_cast<C*>(W.dap.p);
    const int cg_max = std::max(500, op.nn);
    const double pcg_tol = prm.divcorr_pcg_tol > 0 ? prm.divcorr_pcg_tol : 1e-10;
    launch_dc_begin(s, W.st.p, mask);
    for (int it = 0; it < prm.div_corr_iters; ++it) {
        D::divj(s, op.nn, it, prm.div_corr_tol, cg_max, pcg_tol, op.nt_ptr.p, op.nt_edge.p, op.nt_val.p,
                op.ms.p, op.dc_bc_node.p, ldiag, E, phi0, dr, dz, dp, dphi, W.st.p, W.part.p);
        HIPCHK(hipGetLastError());
        std::vector<SysState> h = read_states(op, W.st);
        bool any_dc = false, any_pcg = false;
        for (int i = 0; i < NS; ++i) {
            if (h[i].dc_status == kRunning) any_dc = true;
            if (h[i].pcg_status == kRunning) any_pcg = true;
        }
        if (!any_dc) break;
        int chunk = 8;
        auto pcg_iter = [&] {
            D::lp(s, op.nn, op.l_ptr.p, op.l_colidx.p, lval, dp, dap, W.st.p, W.part.p);
            D::update(s, op.nn, ldiag, dp, dap, dphi, dr, W.st.p, W.part.p);
            D::dir(s, op.nn, ldiag, dr, dp, W.st.p);
        };
        while (any_pcg) {
            if (use_graphs()) {
                const int kind = std::is_same<V, float>::value ? 101 : 100;
                hipGraphExec_t ex = W.graphs.get(s, W.dp.p, kind, kGraphPcgIters, [&] {
                    for (int k = 0; k < kGraphPcgIters; ++k) pcg_iter();
                });
                HIPCHK(hipGraphLaunch(ex, s));
            } else {
                for (int k = 0; k < chunk; ++k) pcg_iter();
                HIPCHK(hipGetLastError());
                chunk = std::min(chunk * 2, 64);
            }
            h = read_states(op, W.st);
            any_pcg = false;
            for (int i = 0; i < NS; ++i)
                if (h[i].pcg_status == kRunning) any_pcg = true;
        }
        D::correct(s, op.npad, op.g_col.p, op.g_val.p, op.dc_bc_edge.p, dphi, E, W.st.p);
        HIPCHK(hipGetLastError());
    }
}

// DivCorr of the masked systems of a width-NA batch at width NB: E (and
// phi0) move into the NB workspace, are corrected there and E moves back.
template <int NA, int NB>
void divcorr_narrow(m3d_operator& op, double2* E, const double2* phi0, unsigned mask,
                    const m3d_solve_params& prm) {
    Work64<NB>& WB = ready_workspace<NB>(op);
    ensure_dc_vectors<NB>(op, WB);
    hipStream_t s = op.ctx->stream;
    std::vector<int> map(NB, -1);
    int j = 0;
    for (int i = 0; i < NA; ++i)
        if (mask & (1u << i)) map[j++] = i;
    HIPCHK(hipMemcpyAsync(WB.flags.p, map.data(), NB * sizeof(int), hipMemcpyHostToDevice, s));
    launch_regroup<double>(s, op.npad, NA, NB, WB.flags.p, E, WB.x.p, true);
    if (phi0)
        launch_regroup<double>(s, op.nn, NA, NB, WB.flags.p, const_cast<double2*>(phi0), WB.phi0.p, true);
    HIPCHK(hipGetLastError());
    const unsigned nmask = (j >= 32) ? 0xFFFFFFFFu : ((1u << j) - 1u);
    divcorr<NB>(op, WB, WB.x.p, phi0 ? WB.phi0.p : nullptr, nmask, prm);
    launch_regroup<double>(s, op.npad, NA, NB, WB.flags.p, E, WB.x.p, false);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
}

// divcorr_poisson_ on the masked systems of the FP64 field E (device order).
// When at most half the batch is masked, the correction runs at the next
// narrower batch width, so its PCG does not carry idle lanes.
template <int NS>
void divcorr(m3d_operator& op, Work64<NS>& W, double2* E, const double2* phi0, unsigned mask,
             const m3d_solve_params& prm) {
    if (!op.dc_set || mask == 0) return;
    const int k = __builtin_popcount(mask);
    if (NS > 1 && use_compaction() && k <= NS / 2) {
        if (k == 1) divcorr_narrow<NS, 1>(op, E, phi0, mask, prm);
        else if (k == 2) divcorr_narrow<NS, 2>(op, E, phi0, mask, prm);
        else if (k <= 4) divcorr_narrow<NS, 4>(op, E, phi0, mask, prm);
        else if (k <= 8) divcorr_narrow<NS, 8>(op, E, phi0, mask, prm);
        else divcorr_narrow<NS, 16>(op, E, phi0, mask, prm);
        return;
    }
    ensure_dc_vectors<NS>(op, W);
    if (dc_in_fp32(prm)) {
        ensure_laplacian32(op);
        divcorr_pcg<NS, float>(op, W, E, phi0, mask, prm, op.l_vals32.p, op.l_diag32.p);
    } else {
        divcorr_pcg<NS, double>(op, W, E, phi0, mask, prm, op.l_vals.p, op.l_diag.p);
    }
}

template <int NS>
std::vector<double> true_residual(m3d_operator& op, Work64<NS>& W, const double2* x,
                                  const double2* b, double2* r) {
    Launch64<NS>::true_res(op.ctx->stream, op.view<double>(), W.omega.p, x, b, r, W.st.p, W.part.p);
    HIPCHK(hipGetLastError());
    std::vector<SysState> h = read_states(op, W.st);
    std::vector<double> out(NS);
    for (int i = 0; i < NS; ++i) out[i] = h[i].true_res;
    return out;
}

template <int NS>
void interp_all(m3d_operator& op, double2* x) {
    for (size_t l = 0; l + 1 < op.hang_level_start.size(); ++l) {
        const int a = op.hang_level_start[l], b = op.hang_level_start[l + 1];
        Launch<NS, double>::interp_hanging(op.ctx->stream, b - a, op.hang.p + 3 * size_t(a), x);
    }
}

void fill_result(m3d_solve_result& r, int iters, double res, int status, int info) {
    r.iterations = iters;
    r.residual = res;
    r.converged = (status == kConverged || status == kConvergedEarly) ? 1 : 0;
    r.info = info;
}

// FP64 BiCGStab, or complex64 BiCGStab inside FP64 iterative refinement, on
// the FP64 system vectors W.b / W.x.  Returns per-system outcome.
template <int NS>
BicgOutcome solve_linear(m3d_operator& op, Work64<NS>& W, const double2* b, double2* x,
                         unsigned mask, int maxiter, const m3d_solve_params& prm, bool x_zero,
                         int* outer_steps) {
    if (prm.precision != M3D_PREC_MIXED) {
        auto& w = vec64<NS>(op, W);
        *outer_steps = 1;
        return bicgstab<NS, double>(op, W, w, b, x, mask, maxiter, prm.tol, x_zero);
    }
    // Iterative refinement: r = b - A x in FP64, A d = r in complex64 to a
    // loose relative tolerance with a capped iteration count, x += d; a
    // system converges when the FP64 true residual ||b - A x|| / ||b|| < tol.
    // Refinement contracts only while kappa(A) * eps32 < 1; a system whose
    // true residual falls by less than 2x in a step (the curl-curl
    // near-null space at low frequency) is finished by FP64 BiCGStab from the
    // refined x, so the result always meets the FP64 criterion.
    auto& w = vec32<NS>(op, W);
    hipStream_t s = op.ctx->stream;
    const size_t nv = size_t(op.npad) * NS;
    if (x_zero) Launch<NS, double>::fill_zero(s, nv, x);
    BicgOutcome total{};
    unsigned active = mask, fallback = 0;
    const double inner_tol = prm.inner_tol > 0 ? prm.inner_tol : 1e-3;
    const int inner_cap = std::max(50, std::min(maxiter, 400));
    int steps = 0;
    std::vector<double> scale(NS, 1.0), prev(NS, 1e300);
    std::vector<int> used(NS, 0);
    for (int step = 0; step < 40 && active; ++step) {
        std::vector<double> tr = true_residual<NS>(op, W, x, b, W.rres.p);
        std::vector<SysState> h = read_states(op, W.st);
        for (int i = 0; i < NS; ++i) {
            if (!(active & (1u << i))) continue;
            total.residual[i] = tr[i];
            if (tr[i] < prm.tol) {
                active &= ~(1u << i);
                total.status[i] = kConverged;
                total.info[i] = 0;
            } else if (tr[i] > 0.5 * prev[i] || used[i] >= maxiter) {
                active &= ~(1u << i);
                fallback |= 1u << i;
            }
            prev[i] = tr[i];
        }
        if (!active) break;
        // Scale each correction system to a unit right-hand side before FP32.
        for (int i = 0; i < NS; ++i) {
            const double rn = tr[i] * h[i].true_bnorm;
            scale[i] = rn > 0 ? rn : 1.0;
        }
        std::vector<double> inv(NS);
        for (int i = 0; i < NS; ++i) inv[i] = 1.0 / scale[i];
        HIPCHK(hipMemcpyAsync(W.scale.p, inv.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
        Launch64<NS>::demote(s, nv, W.rres.p, W.scale.p, w.b.p);
        Launch<NS, float>::fill_zero(s, nv, w.x.p);
        BicgOutcome o = bicgstab<NS, float>(op, W, w, w.b.p, w.x.p, active, inner_cap, inner_tol, true);
        HIPCHK(hipMemcpyAsync(W.scale.p, scale.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
        Launch64<NS>::add_scaled(s, nv, w.x.p, W.scale.p, x);
        for (int i = 0; i < NS; ++i)
            if (active & (1u << i)) {
                total.iterations[i] += o.iterations[i];
                used[i] += o.iterations[i];
            }
        ++steps;
    }
    fallback |= active;  // out of refinement steps
    if (fallback) {
        auto& w64 = vec64<NS>(op, W);
        int left = 0;
        for (int i = 0; i < NS; ++i)
            if (fallback & (1u << i)) left = std::max(left, maxiter - used[i]);
        BicgOutcome o = bicgstab<NS, double>(op, W, w64, b, x, fallback, std::max(left, 1), prm.tol, false);
        for (int i = 0; i < NS; ++i) {
            if (!(fallback & (1u << i))) continue;
            total.iterations[i] += o.iterations[i];
            total.status[i] = o.status[i];
            total.info[i] = o.info[i];
            total.residual[i] = o.residual[i];
        }
    }
    *outer_steps = steps + (fallback ? 1 : 0);
    return total;
}

// Mixed precision only.  Refinement hands a system whose FP64 true residual
// stops falling to FP64 BiCGStab from the refined x; at low frequency that
// finish can still diverge or break down, where the same BiCGStab from the
// original start converges (which trajectory happens depends on rounding).
// Systems of `mask` that end unconverged with a true residual above tol are
// therefore solved again by the reference's FP64 BiCGStab from their start:
// zero, or the caller's system-major guess `caller_x` (packed again).  So the
// mixed path never ends worse than the FP64 method from the same start.  The
// repeat's outcome replaces the mixed one; iterations add up.
template <int NS>
void restart_failed(m3d_operator& op, Work64<NS>& W, const double2* b, double2* x, unsigned mask,
                    int nsys, const double2* caller_x, const m3d_solve_params& prm, BicgOutcome& o) {
    if (prm.precision != M3D_PREC_MIXED) return;
    unsigned cand = 0;
    for (int i = 0; i < nsys; ++i)
        if ((mask & (1u << i)) && o.status[i] != kConverged && o.status[i] != kConvergedEarly) cand |= 1u << i;
    if (!cand) return;
    std::vector<double> tr = true_residual<NS>(op, W, x, b, nullptr);
    unsigned failed = 0;
    for (int i = 0; i < nsys; ++i)
        if ((cand & (1u << i)) && !(tr[i] < prm.tol)) failed |= 1u << i;
    if (!failed) return;
    hipStream_t s = op.ctx->stream;
    const double2* x0 = nullptr;
    if (caller_x) {
        Launch<NS, double>::pack(s, op.npad, op.n, nsys, op.d_perm.p, caller_x, W.rres.p);
        x0 = W.rres.p;
    }
    std::vector<int> flag(NS, 0);
    for (int i = 0; i < NS; ++i) flag[i] = int((failed >> i) & 1u);
    HIPCHK(hipMemcpyAsync(W.flags.p, flag.data(), NS * sizeof(int), hipMemcpyHostToDevice, s));
    Launch64<NS>::copy_systems(s, op.npad, W.flags.p, x0, x);
    HIPCHK(hipGetLastError());
    auto& w64 = vec64<NS>(op, W);
    BicgOutcome r = bicgstab<NS, double>(op, W, w64, b, x, failed, prm.maxiter, prm.tol, x0 == nullptr);
    for (int i = 0; i < nsys; ++i) {
        if (!(failed & (1u << i))) continue;
        o.iterations[i] += r.iterations[i];
        o.residual[i] = r.residual[i];
        o.status[i] = r.status[i];
        o.info[i] = r.info[i];
    }
}

// ModEM-style interleaved BiCGStab + DivCorr on A x = b from the current x
// (solve_with_divcorr_, forward_solver_fv.cpp:778-888): phi0 = G^T b / (i w),
// an initial DivCorr, then bursts of divcorr_iter_per_dc BiCGStab iterations,
// each followed by DivCorr and an FP64 true residual.  A system stops when
// the burst or the true residual reaches tol, when a burst ends above
// residual 1 (x zeroed, diverged), or when a burst gains less than 5 %
// (accepted as is).  Iterations accumulate into res.  With `unfinished`
// non-null, systems that diverged, stagnated or used every burst are not
// accepted but reported there (their x is left as it is; zero if diverged).
template <int NS>
void interleaved(m3d_operator& op, Work64<NS>& W, const double2* b, double2* x, unsigned mask, int nsys,
                 const m3d_solve_params& prm, m3d_solve_result* res, unsigned* unfinished) {
    using L64 = Launch64<NS>;
    hipStream_t s = op.ctx->stream;
    ensure_dc_vectors<NS>(op, W);
    L64::phi0(s, op.nn, op.nt_ptr.p, op.nt_edge.p, op.nt_val.p, op.dc_dead.p, W.omega.p, b, W.phi0.p);
    divcorr<NS>(op, W, x, W.phi0.p, mask, prm);
    unsigned active = mask, left = 0;
    std::vector<double> prev(NS, 1e30);
    std::vector<int> zero(NS, 0);
    for (int outer = 0; outer < prm.divcorr_outer_max && active; ++outer) {
        int steps = 1;
        BicgOutcome o = solve_linear<NS>(op, W, b, x, active, prm.divcorr_iter_per_dc, prm, false, &steps);
        for (int i = 0; i < nsys; ++i) {
            if (!(active & (1u << i))) continue;
            res[i].iterations += o.iterations[i];
            res[i].residual = o.residual[i];
            res[i].info = o.info[i];
            res[i].outer = outer + 1;
            if (o.residual[i] > 1.0) {
                std::fill(zero.begin(), zero.end(), 0);
                zero[i] = 1;
                res[i].diverged = unfinished ? 0 : 1;
                active &= ~(1u << i);
                left |= 1u << i;
                HIPCHK(hipMemcpyAsync(W.flags.p, zero.data(), NS * sizeof(int), hipMemcpyHostToDevice, s));
                L64::zero_systems(s, op.npad, W.flags.p, x);
                HIPCHK(hipStreamSynchronize(s));
            } else if (o.status[i] == kConverged || o.status[i] == kConvergedEarly) {
                res[i].converged = 1;
                active &= ~(1u << i);
            }
        }
        if (!active) break;
        divcorr<NS>(op, W, x, W.phi0.p, active, prm);
        std::vector<double> tr = true_residual<NS>(op, W, x, b, nullptr);
        for (int i = 0; i < nsys; ++i) {
            if (!(active & (1u << i))) continue;
            if (tr[i] < prm.tol) {
                res[i].converged = 1;
                active &= ~(1u << i);
            } else if (outer > 0 && tr[i] > 0.95 * prev[i]) {
                active &= ~(1u << i);  // stagnation: accept (:866-873)
                left |= 1u << i;
                if (!unfinished) res[i].stagnated = 1;
            } else {
                prev[i] = tr[i];
            }
        }
    }
    left |= active;  // out of bursts
    if (unfinished) *unfinished = left;
}

template <int NS>
void solve_impl(m3d_operator& op, int nsys, int mode, const double* omega, const double2* d_rhs,
                const double2* d_aux, double2* d_x, const m3d_solve_params& prm,
                m3d_solve_result* res) {
    Work64<NS>& W = ready_workspace<NS>(op);
    hipStream_t s = op.ctx->stream;
    std::vector<double> om(NS, 0.0);
    for (int i = 0; i < nsys; ++i) om[i] = omega[i];
    W.h_omega = om;
    HIPCHK(hipMemcpyAsync(W.omega.p, om.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
    const unsigned mask = (nsys >= 32) ? 0xFFFFFFFFu : ((1u << nsys) - 1u);
    using L = Launch<NS, double>;
    using L64 = Launch64<NS>;
    const size_t nv = size_t(op.npad) * NS;
    L::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_rhs, W.b.p);
    for (int i = 0; i < nsys; ++i) res[i] = m3d_solve_result{};

    const bool dc = prm.div_correction && op.dc_set;
    if (mode == M3D_MODE_FORWARD) {
        L::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_aux, W.aux.p);
        L::spmv(s, op.view<double>(), W.omega.p, W.aux.p, W.b.p, W.tmp.p);  // rhs_sec
        bool x_zero = true;
        std::vector<int> used(NS, 0);
        std::vector<double> guess_res(NS, 0.0);
        if (prm.use_initial_guess) {
            // Caller's E_sec guess (modes 2 and 3: scaled by the least-squares
            // factor first), kept only where it lowers the residual below
            // ||b|| (mode 2: 0.5 ||b||; a guess that removes less than half the
            // residual often adds rough error that BiCGStab must undo);
            // elsewhere the reference's zero start.
            L::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_x, W.x.p);
            L64::fix_constrained(s, op.npad, op.constrained.p, W.tmp.p, W.x.p);
            const bool scaled = prm.use_initial_guess >= 2;
            if (scaled) {
                L64::scale_guess(s, op.view<double>(), W.omega.p, W.tmp.p, W.x.p, W.st.p, W.part.p);
                L64::fix_constrained(s, op.npad, op.constrained.p, W.tmp.p, W.x.p);
            }
            std::vector<double> tr = true_residual<NS>(op, W, W.x.p, W.tmp.p, nullptr);
            std::vector<int> zero(NS, 0);
            const double keep = prm.use_initial_guess == 2 ? 0.5 : 1.0;
            for (int i = 0; i < NS; ++i) {
                used[i] = (i < nsys && tr[i] < keep) ? 1 : 0;
                zero[i] = used[i] ? 0 : 1;
                if (i < nsys) guess_res[i] = tr[i];
            }
            HIPCHK(hipMemcpyAsync(W.flags.p, zero.data(), NS * sizeof(int), hipMemcpyHostToDevice, s));
            L64::zero_systems(s, op.npad, W.flags.p, W.x.p);
            x_zero = false;
        } else {
            L::fill_zero(s, nv, W.x.p);
        }
        bool all_omega = true;
        for (int i = 0; i < nsys; ++i)
            if (!(omega[i] > 0)) all_omega = false;
        if (prm.interleaved_forward && dc && all_omega) {
            // E_sec by interleaved BiCGStab + DivCorr bursts; systems that do
            // not converge that way finish with the reference's full BiCGStab
            // from where they stand.
            unsigned left = 0;
            interleaved<NS>(op, W, W.tmp.p, W.x.p, mask, nsys, prm, res, &left);
            if (left) {
                int steps = 1;
                BicgOutcome o = solve_linear<NS>(op, W, W.tmp.p, W.x.p, left, prm.maxiter, prm, false, &steps);
                restart_failed<NS>(op, W, W.tmp.p, W.x.p, left, nsys, nullptr, prm, o);
                for (int i = 0; i < nsys; ++i) {
                    if (!(left & (1u << i))) continue;
                    const int before = res[i].iterations;
                    fill_result(res[i], o.iterations[i], o.residual[i], o.status[i], o.info[i]);
                    res[i].iterations += before;
                    res[i].outer += steps;
                }
            }
        } else {
            int steps = 1;
            BicgOutcome o = solve_linear<NS>(op, W, W.tmp.p, W.x.p, mask, prm.maxiter, prm, x_zero, &steps);
            restart_failed<NS>(op, W, W.tmp.p, W.x.p, mask, nsys, nullptr, prm, o);
            for (int i = 0; i < nsys; ++i) {
                fill_result(res[i], o.iterations[i], o.residual[i], o.status[i], o.info[i]);
                res[i].outer = steps;
            }
        }
        interp_all<NS>(op, W.x.p);
        L64::forward_total(s, op.npad, op.bc.p, W.aux.p, W.b.p, W.x.p);
        if (dc) {
            // The post-hoc correction of the total field is the reference's
            // (PCG to 1e-10) whatever the bursts used.
            m3d_solve_params post = prm;
            post.divcorr_pcg_tol = 0.0;
            divcorr<NS>(op, W, W.x.p, nullptr, mask, post);
        }
        for (int i = 0; i < nsys; ++i) {
            res[i].guess_used = used[i];
            res[i].guess_residual = guess_res[i];
        }
    } else if (mode == M3D_MODE_INNER || mode == M3D_MODE_SYSTEM) {
        L::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_x, W.x.p);
        int steps = 1;
        BicgOutcome o = solve_linear<NS>(op, W, W.b.p, W.x.p, mask, prm.maxiter, prm, false, &steps);
        restart_failed<NS>(op, W, W.b.p, W.x.p, mask, nsys, d_x, prm, o);
        if (mode == M3D_MODE_INNER) interp_all<NS>(op, W.x.p);
        std::vector<int> zero(NS, 0);
        bool any = false;
        for (int i = 0; i < nsys; ++i) {
            fill_result(res[i], o.iterations[i], o.residual[i], o.status[i], o.info[i]);
            res[i].outer = steps;
            if (mode == M3D_MODE_SYSTEM && !res[i].converged && res[i].residual > 1.0) {
                zero[i] = 1;
                res[i].diverged = 1;
                any = true;
            }
        }
        if (any) {
            HIPCHK(hipMemcpyAsync(W.flags.p, zero.data(), NS * sizeof(int), hipMemcpyHostToDevice, s));
            L64::zero_systems(s, op.npad, W.flags.p, W.x.p);
        }
    } else if (mode == M3D_MODE_INTERLEAVED) {
        L::fill_zero(s, nv, W.x.p);
        bool use_dc = dc;
        for (int i = 0; i < nsys; ++i)
            if (!(omega[i] > 0)) use_dc = false;
        if (!use_dc) {
            // solve_system_ path (forward_solver_fv.cpp:921-923).
            int steps = 1;
            BicgOutcome o = solve_linear<NS>(op, W, W.b.p, W.x.p, mask, prm.maxiter, prm, true, &steps);
            restart_failed<NS>(op, W, W.b.p, W.x.p, mask, nsys, nullptr, prm, o);
            std::vector<int> zero(NS, 0);
            bool any = false;
            for (int i = 0; i < nsys; ++i) {
                fill_result(res[i], o.iterations[i], o.residual[i], o.status[i], o.info[i]);
                res[i].outer = steps;
                if (!res[i].converged && res[i].residual > 1.0) {
                    zero[i] = 1; res[i].diverged = 1; any = true;
                }
            }
            if (any) {
                HIPCHK(hipMemcpyAsync(W.flags.p, zero.data(), NS * sizeof(int), hipMemcpyHostToDevice, s));
                L64::zero_systems(s, op.npad, W.flags.p, W.x.p);
            }
        } else {
            interleaved<NS>(op, W, W.b.p, W.x.p, mask, nsys, prm, res, nullptr);
        }
    } else {
        throw Fail(M3D_ERR_ARG, "unknown solve mode");
    }
    std::vector<double> tr = true_residual<NS>(op, W, W.x.p, W.b.p, nullptr);
    for (int i = 0; i < nsys; ++i) res[i].true_residual = tr[i];
    L::unpack(s, op.npad, op.n, nsys, op.d_perm.p, W.x.p, d_x);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
}

void check_ready(const m3d_operator* op) {
    if (!op) throw Fail(M3D_ERR_ARG, "null operator");
    if (!op->mass_set) throw Fail(M3D_ERR_STATE, "m3d_operator_set_mass has not been called");
}

void check_params(const m3d_solve_params* p) {
    if (!p) throw Fail(M3D_ERR_ARG, "null params");
    if (!(p->tol > 0) || p->maxiter <= 0) throw Fail(M3D_ERR_ARG, "tol and maxiter must be positive");
    if (p->div_correction && (p->div_corr_iters < 0 || p->divcorr_outer_max < 0 || p->divcorr_iter_per_dc <= 0))
        throw Fail(M3D_ERR_ARG, "invalid DivCorr parameters");
    if (p->precision != M3D_PREC_FP64 && p->precision != M3D_PREC_MIXED)
        throw Fail(M3D_ERR_ARG, "precision must be M3D_PREC_FP64 or M3D_PREC_MIXED");
    if (p->use_initial_guess < 0 || p->use_initial_guess > 3)
        throw Fail(M3D_ERR_ARG, "use_initial_guess must be 0, 1, 2 or 3");
    if (p->interleaved_forward != 0 && p->interleaved_forward != 1)
        throw Fail(M3D_ERR_ARG, "interleaved_forward must be 0 or 1");
    if (!(p->divcorr_pcg_tol >= 0.0 && p->divcorr_pcg_tol < 1.0))
        throw Fail(M3D_ERR_ARG, "divcorr_pcg_tol must be in [0, 1)");
}

// Dispatch a batch of any size onto the fixed batch widths (<= 32 per launch).
void solve_dispatch(m3d_operator* op, int nsys, int mode, const double* omega, const double2* d_rhs,
                    const double2* d_aux, double2* d_x, const m3d_solve_params& prm,
                    m3d_solve_result* res) {
    for (int s0 = 0; s0 < nsys; s0 += kMaxSys) {
        const int k = std::min(kMaxSys, nsys - s0);
        const size_t off = size_t(s0) * op->n;
        const double2* rb = d_rhs + off;
        const double2* ab = d_aux ? d_aux + off : nullptr;
        double2* xb = d_x + off;
        if (k == 1) solve_impl<1>(*op, k, mode, omega + s0, rb, ab, xb, prm, res + s0);
        else if (k == 2) solve_impl<2>(*op, k, mode, omega + s0, rb, ab, xb, prm, res + s0);
        else if (k <= 4) solve_impl<4>(*op, k, mode, omega + s0, rb, ab, xb, prm, res + s0);
        else if (k <= 8) solve_impl<8>(*op, k, mode, omega + s0, rb, ab, xb, prm, res + s0);
        else if (k <= 16) solve_impl<16>(*op, k, mode, omega + s0, rb, ab, xb, prm, res + s0);
        else solve_impl<32>(*op, k, mode, omega + s0, rb, ab, xb, prm, res + s0);
    }
}

// Greedy first-fit colouring of the coupling graph in caller order.
// Greedy colouring of the rows visited in `order` (smallest colour not used
// by an already coloured neighbour); constrained and empty rows get colour 0.
int greedy_colours(int n, const int* ptr, const int* col, const unsigned char* constrained,
                   const std::vector<int>& order, std::vector<int>& c) {
    c.assign(n, -1);
    std::vector<int> stamp(64, -1);
    int nc = 1;
    for (int i : order) {
        if (constrained[i] || ptr[i + 1] == ptr[i]) { c[i] = 0; continue; }
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const int cj = c[col[k]];
            if (cj >= 0) {
                if (cj >= int(stamp.size())) stamp.resize(cj + 1, -1);
                stamp[cj] = i;
            }
        }
        int pick = 0;
        while (pick < int(stamp.size()) && stamp[pick] == i) ++pick;
        if (pick >= int(stamp.size())) stamp.resize(pick + 1, -1);
        c[i] = pick;
        nc = std::max(nc, pick + 1);
    }
    return nc;
}

// Multicolouring for the SSOR sweeps: greedy in row order, then iterated
// greedy passes (rows re-visited class by class, highest colour first), which
// never add a colour and usually remove the sparse last ones: every colour is
// one kernel launch per sweep.
std::vector<int> colour_rows(int n, const int* ptr, const int* col, const unsigned char* constrained,
                             int* ncolours) {
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::vector<int> c, trial;
    int nc = greedy_colours(n, ptr, col, constrained, order, c);
    for (int pass = 0; pass < 4; ++pass) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return c[a] > c[b]; });
        const int nt = greedy_colours(n, ptr, col, constrained, order, trial);
        if (nt > nc) break;
        c.swap(trial);
        nc = nt;
        for (int i = 0; i < n; ++i) order[i] = i;
    }
    *ncolours = nc;
    return c;
}

LaunchShape default_shape(int width) {
    LaunchShape sh;
    sh.sweep_sub = 1;  // with grouped gathers one lane per row keeps enough loads in flight
    sh.sweep_cap = 4096;
    sh.row_cap = kMaxBlocks;
    sh.sweep_chunk = 1;
    // Packed blocks: fewer bytes per entry shorten every gather (one
    // BiCGStab iteration on a 1.45 M-edge mesh -12 % at 1 system, -5 % at 2,
    // -3 to -14 % at 8-32 with the row's lanes sharing the couplings).
    sh.packed = 1;
    // Couplings shared among a row's lanes: the 16-system SpMV 0.58 -> 0.45 ms,
    // an FP64 BiCGStab iteration -3 % at 8 systems and -5 % at 16, a complex64
    // one -3 % and -8 % on the same mesh.
    sh.wave = width >= 8 ? 1 : 0;
    // The plain product over LDS-staged row tiles where the operator has them.
    sh.tiled = width >= 8 ? 1 : 0;
    return sh;
}

// Row tiles of the LDS-staged plain product at batch width 1 << wi (the
// layout DevTiles describes): runs of rows in the caller's edge order, then
// the padding rows, each tile closed before the distinct rows it stages (its
// rows and the rows their couplings read) would exceed max_stage, or when it
// holds max_slices slices of H rows (one per wavefront of the workgroup).
void build_tiles(m3d_operator& op, const RowEntries& E, int wi, int max_stage, int max_slices, hipStream_t s) {
    const int H = kSlice >> wi;
    const size_t max_rows = size_t(max_slices) * H;
    int lbits = 1;
    while ((1 << lbits) < max_stage) ++lbits;
    std::vector<int> stage_off{0}, stage_rows, slice_off{0};
    std::vector<int4> slc;
    std::vector<int2> slot;
    std::vector<unsigned> words;
    std::vector<int> mark(op.npad, -1), local(op.npad, 0);
    std::vector<int> rows, st;  // the open tile's rows and staged rows
    int tile = 0, most = 0;
    auto close = [&]() {
        if (rows.empty()) return;
        std::sort(st.begin(), st.end());
        for (size_t i = 0; i < st.size(); ++i) local[st[i]] = int(i);
        stage_rows.insert(stage_rows.end(), st.begin(), st.end());
        stage_off.push_back(int(stage_rows.size()));
        most = std::max(most, int(st.size()));
        // Longest rows first, so the rows of a slice have similar widths.
        auto len = [&](int r) { return E.ptr[r + 1] - E.ptr[r]; };
        std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return len(a) > len(b); });
        for (size_t r0 = 0; r0 < rows.size(); r0 += H) {
            const size_t r1 = std::min(rows.size(), r0 + H);
            int W = 0;
            for (size_t q = r0; q < r1; ++q) W = std::max(W, len(rows[q]));
            const size_t off = words.size();
            words.resize(off + size_t(W) * H, 0u);
            slc.push_back(make_int4(int(off), W, int(slot.size()), int(r1 - r0)));
            for (int l = 0; l < H; ++l) {
                if (r0 + l >= r1) {
                    slot.push_back(make_int2(-1, 0));
                    continue;
                }
                const int r = rows[r0 + l];
                slot.push_back(make_int2(r, local[r]));
                for (int k = E.ptr[r]; k < E.ptr[r + 1]; ++k)
                    words[off + size_t(k - E.ptr[r]) * H + l] =
                        unsigned(local[E.col[k]]) | (unsigned(E.tab[k]) << lbits);
            }
        }
        slice_off.push_back(int(slc.size()));
        ++tile;
        rows.clear();
        st.clear();
    };
    auto add_row = [&](int r) {
        int fresh = mark[r] != tile ? 1 : 0;
        for (int k = E.ptr[r]; k < E.ptr[r + 1]; ++k) fresh += mark[E.col[k]] != tile ? 1 : 0;
        if (!rows.empty() && (int(st.size()) + fresh > max_stage || rows.size() == max_rows)) close();
        auto stage = [&](int j) {
            if (mark[j] != tile) {
                mark[j] = tile;
                st.push_back(j);
            }
        };
        stage(r);
        for (int k = E.ptr[r]; k < E.ptr[r + 1]; ++k) stage(E.col[k]);
        rows.push_back(r);
    };
    for (int e = 0; e < op.n; ++e) add_row(op.iperm[e]);
    close();
    for (int r = 0; r < op.npad; ++r)
        if (op.perm[r] < 0) add_row(r);
    close();
    if (words.size() > size_t(std::numeric_limits<int>::max()))
        throw Fail(M3D_ERR_ARG, "operator too large for int32 tile offsets");
    auto& T = op.tiles[wi];
    T.stage_off.upload(stage_off, s);
    T.stage_rows.upload(stage_rows, s);
    T.slice_off.upload(slice_off, s);
    T.slc.upload(slc, s);
    T.slot.upload(slot, s);
    T.words.upload(words, s);
    T.ntiles = tile;
    T.lbits = lbits;
    T.max_stage = most;
    HIPCHK(hipStreamSynchronize(s));
}

void build_operator(m3d_operator& op, const m3d_operator_desc& d) {
    const int n = d.n;
    hipStream_t s = op.ctx->stream;
    op.n = n;
    for (int p = 0; p < 2; ++p)
        for (int i = 0; i < kWidths; ++i) op.shape[p][i] = default_shape(1 << i);
    op.h_constrained.assign(d.constrained, d.constrained + n);
    int K = 0;
    std::vector<int> colour = colour_rows(n, d.rowptr, d.colidx, d.constrained, &K);
    op.ncolours = K;
    std::vector<int> count(K, 0);
    for (int i = 0; i < n; ++i) count[colour[i]]++;
    op.colour_start.assign(K + 1, 0);
    for (int c = 0; c < K; ++c) op.colour_start[c + 1] = op.colour_start[c] + ((count[c] + kSlice - 1) / kSlice) * kSlice;
    op.npad = op.colour_start[K];
    op.perm.assign(op.npad, -1);
    op.iperm.assign(n, -1);
    // Rows of a colour never couple to each other, so their order inside the
    // colour leaves the multicolour SSOR unchanged.  By default they keep the
    // edge (Morton) order, which keeps the x gathers of neighbouring rows on
    // shared cache lines.  With M3D_ROW_SORT=1 each window of kSortWindow rows
    // is sorted by (lower, upper) coupling counts instead, so the rows of a
    // 64-row slice have similar widths: SELL padding drops from ~35 % to ~7 %
    // on octree meshes, which speeds up wide batches (16-system SpMV -12 %)
    // but slows narrow ones (2-system sweeps +5-10 %), whose 32-byte gathers
    // depend on that locality.
    const int kSortWindow = row_sort() ? 1024 : 1;
    std::vector<int> key(n, 0);
    for (int i = 0; i < n; ++i) {
        int lo = 0, up = 0;
        for (int k = d.rowptr[i]; k < d.rowptr[i + 1]; ++k) {
            const int cj = colour[d.colidx[k]];
            if (cj < colour[i]) ++lo;
            else if (cj > colour[i]) ++up;
        }
        key[i] = lo * 1024 + up;
    }
    std::vector<std::vector<int>> rows_of(K);
    for (int c = 0; c < K; ++c) rows_of[c].reserve(count[c]);
    for (int i = 0; i < n; ++i) rows_of[colour[i]].push_back(i);
    for (int c = 0; c < K; ++c) {
        auto& v = rows_of[c];
        for (size_t a = 0; a < v.size(); a += kSortWindow) {
            const size_t b = std::min(v.size(), a + kSortWindow);
            std::stable_sort(v.begin() + a, v.begin() + b, [&](int x, int y) { return key[x] > key[y]; });
        }
        for (size_t j = 0; j < v.size(); ++j) {
            const int r = op.colour_start[c] + int(j);
            op.perm[r] = v[j];
            op.iperm[v[j]] = r;
        }
    }
    // SELL-64 lower / upper blocks.
    const int nsl = op.npad / kSlice;
    std::vector<int> loff(nsl), lw(nsl), uoff(nsl), uw(nsl);
    long long ltot = 0, utot = 0;
    for (int sl = 0; sl < nsl; ++sl) {
        int ml = 0, mu = 0;
        for (int l = 0; l < kSlice; ++l) {
            const int e = op.perm[sl * kSlice + l];
            if (e < 0) continue;
            int cl = 0, cu = 0;
            for (int k = d.rowptr[e]; k < d.rowptr[e + 1]; ++k)
                (op.iperm[d.colidx[k]] < sl * kSlice + l ? cl : cu)++;
            ml = std::max(ml, cl);
            mu = std::max(mu, cu);
        }
        loff[sl] = int(ltot); lw[sl] = ml; ltot += (long long)ml * kSlice;
        uoff[sl] = int(utot); uw[sl] = mu; utot += (long long)mu * kSlice;
    }
    if (ltot > 2147483647LL || utot > 2147483647LL) throw Fail(M3D_ERR_ARG, "operator too large for int32 offsets");
    std::vector<double> lval(ltot, 0.0), uval(utot, 0.0);
    std::vector<int> lcol(ltot), ucol(utot);
    long long nz = 0;
    for (int sl = 0; sl < nsl; ++sl) {
        for (int l = 0; l < kSlice; ++l) {
            const int row = sl * kSlice + l;
            for (int k = 0; k < lw[sl]; ++k) lcol[loff[sl] + k * kSlice + l] = row;
            for (int k = 0; k < uw[sl]; ++k) ucol[uoff[sl] + k * kSlice + l] = row;
            const int e = op.perm[row];
            if (e < 0) continue;
            int kl = 0, ku = 0;
            for (int k = d.rowptr[e]; k < d.rowptr[e + 1]; ++k) {
                const int j = op.iperm[d.colidx[k]];
                if (j < row) {
                    lval[loff[sl] + kl * kSlice + l] = d.values[k];
                    lcol[loff[sl] + kl * kSlice + l] = j;
                    ++kl;
                } else {
                    uval[uoff[sl] + ku * kSlice + l] = d.values[k];
                    ucol[uoff[sl] + ku * kSlice + l] = j;
                    ++ku;
                }
                ++nz;
            }
        }
    }
    op.stored = ltot + utot;
    op.last_upper_empty = K > 0 && fuse_last_colour();
    for (int sl = K > 0 ? op.colour_start[K - 1] / kSlice : nsl; sl < nsl; ++sl)
        if (uw[sl] != 0) op.last_upper_empty = false;
    op.nnz = nz;
    std::vector<double> ds(op.npad, 1.0);
    std::vector<unsigned char> bc(op.npad, 1), cons(op.npad, 1);
    for (int r = 0; r < op.npad; ++r) {
        const int e = op.perm[r];
        if (e < 0) continue;
        ds[r] = d.constrained[e] ? 1.0 : d.diag_s[e];
        bc[r] = d.bc[e];
        cons[r] = d.constrained[e];
    }
    op.d_perm.upload(op.perm, s);
    std::vector<int4> slc(nsl);
    for (int sl = 0; sl < nsl; ++sl) slc[sl] = make_int4(loff[sl], lw[sl], uoff[sl], uw[sl]);
    op.slc.upload(slc, s);
    // Device rows come in colour order, so the rows that gather one edge's
    // values sit in as many distant row ranges as it has neighbour colours.
    // The caller's edge numbering follows the mesh (Morton order of cells):
    // row blocks sorted by their smallest caller edge bring those rows
    // together for the kernels that may visit blocks in any order.
    std::vector<int> order[kWidths];
    if (spmv_order()) {
        for (int wi = 0; wi < kWidths; ++wi) {
            const int rpb = kBlock >> wi;  // widths 1, 2, 4, ..., 32
            const int nrb = (op.npad + rpb - 1) / rpb;
            std::vector<std::pair<int, int>> key(nrb);
            for (int b = 0; b < nrb; ++b) {
                int k = std::numeric_limits<int>::max();
                for (int r = b * rpb; r < std::min(op.npad, (b + 1) * rpb); ++r)
                    if (op.perm[r] >= 0) k = std::min(k, op.perm[r]);
                key[b] = {k, b};
            }
            std::sort(key.begin(), key.end());
            order[wi].resize(nrb);
            for (int b = 0; b < nrb; ++b) order[wi][b] = key[b].second;
            op.order[wi].upload(order[wi], s);
        }
    }
    // The couplings of an octree mesh take few distinct values (cell sizes
    // are powers of two: 82 values among 15 M entries on a five-level mesh),
    // so the blocks are stored as one word per entry, column | value index <<
    // cbits, with a value table, when the table index fits in the bits the
    // columns leave and the table stays small enough to sit in cache.  The
    // values and their order of summation are unchanged.
    std::vector<double> tab{0.0};
    std::vector<float> tab32, lval32, uval32;
    std::vector<unsigned> lpk, upk;
    bool pack = packed_sell();
    int cbits = 1;
    while ((1LL << cbits) < op.npad) ++cbits;
    if (pack) {
        FlatIndex tix;
        tix.insert(bits_of(0.0), 0);
        auto scan = [&](const std::vector<double>& v) {
            for (double a : v) {
                if (!tix.insert(bits_of(a), int(tab.size()))) continue;
                tab.push_back(a);
                if (tab.size() > size_t(kMaxValueTable)) return false;
            }
            return true;
        };
        pack = scan(lval) && scan(uval);
        int vbits = 1;
        while ((1LL << vbits) < (long long)tab.size()) ++vbits;
        pack = pack && cbits + vbits <= 32;
        if (pack) {
            auto words = [&](const std::vector<double>& v, const std::vector<int>& c, std::vector<unsigned>& out) {
                out.resize(v.size());
                for (size_t i = 0; i < v.size(); ++i)
                    out[i] = unsigned(c[i]) | (unsigned(tix.find(bits_of(v[i]))) << cbits);
            };
            words(lval, lcol, lpk);
            words(uval, ucol, upk);
            tab32.assign(tab.begin(), tab.end());
            op.l_pk.upload(lpk, s);
            op.u_pk.upload(upk, s);
            op.vtab.upload(tab, s);
            op.vtab32.upload(tab32, s);
            op.cbits = cbits;
            op.ntab = int(tab.size());
            if (tiled_spmv()) {
                RowEntries E;
                E.ptr.assign(op.npad + 1, 0);
                E.col.reserve(size_t(nz));
                E.tab.reserve(size_t(nz));
                for (int r = 0; r < op.npad; ++r) {
                    const int e = op.perm[r];
                    for (int upper = 0; e >= 0 && upper < 2; ++upper)
                        for (int k = d.rowptr[e]; k < d.rowptr[e + 1]; ++k) {
                            const int j = op.iperm[d.colidx[k]];
                            if ((j < r) == (upper == 1)) continue;
                            E.col.push_back(j);
                            E.tab.push_back(tix.find(bits_of(d.values[k])));
                        }
                    E.ptr[r + 1] = int(E.col.size());
                }
                for (int wi = width_index(8); wi < kWidths; ++wi)
                    build_tiles(op, E, wi, tile_stage_bytes(op.ntab) / (int(sizeof(double2)) << wi), kTileWaves, s);
            }
        }
    }
    // The separate blocks stay resident beside the packed ones: the launch
    // shape picks one layout per batch width.
    op.l_val.upload(lval, s); op.u_val.upload(uval, s); op.l_col.upload(lcol, s); op.u_col.upload(ucol, s);
    lval32.assign(lval.begin(), lval.end());
    uval32.assign(uval.begin(), uval.end());
    op.l_val32.upload(lval32, s); op.u_val32.upload(uval32, s);
    op.diag_s.upload(ds, s);
    op.bc.upload(bc, s);
    op.constrained.upload(cons, s);
    std::vector<float> ds32(ds.begin(), ds.end());
    op.diag_s32.upload(ds32, s);

    // Hanging pairs in dependency levels that keep the sequential semantics
    // of the caller's list order.
    std::vector<int> lastw(n, -1), maxr(n, -1), level(d.n_hanging, 0);
    int nlev = 0;
    for (int p = 0; p < d.n_hanging; ++p) {
        const int c = d.hanging[3 * p], a = d.hanging[3 * p + 1], b = d.hanging[3 * p + 2];
        if (c < 0 || c >= n || a < 0 || a >= n || b < 0 || b >= n) throw Fail(M3D_ERR_ARG, "hanging index out of range");
        int lv = 0;
        lv = std::max(lv, lastw[a] + 1);
        lv = std::max(lv, lastw[b] + 1);
        lv = std::max(lv, maxr[c] + 1);
        lv = std::max(lv, lastw[c] + 1);
        level[p] = lv;
        maxr[a] = std::max(maxr[a], lv);
        maxr[b] = std::max(maxr[b], lv);
        lastw[c] = lv;
        nlev = std::max(nlev, lv + 1);
    }
    std::vector<int> tri;
    op.hang_level_start.assign(1, 0);
    for (int l = 0; l < nlev; ++l) {
        for (int p = 0; p < d.n_hanging; ++p) {
            if (level[p] != l) continue;
            tri.push_back(op.iperm[d.hanging[3 * p]]);
            tri.push_back(op.iperm[d.hanging[3 * p + 1]]);
            tri.push_back(op.iperm[d.hanging[3 * p + 2]]);
        }
        op.hang_level_start.push_back(int(tri.size() / 3));
    }
    op.hang.upload(tri, s);
    HIPCHK(hipStreamSynchronize(s));
}

// Device bytes held by a set of buffers.
template <typename T>
size_t nbytes(const DBuf<T>& b) {
    return b.p ? b.n * sizeof(T) : 0;
}
template <typename T, typename... R>
size_t nbytes(const DBuf<T>& b, const R&... rest) {
    return nbytes(b) + nbytes(rest...);
}

template <int NS, typename V>
size_t work_bytes(const Work<NS, V>* w) {
    return w ? nbytes(w->b, w->x, w->r, w->r0, w->p, w->v, w->s, w->t, w->y, w->ph, w->sh) : 0;
}
template <int NS>
size_t work64_bytes(const Work64<NS>* W) {
    if (!W) return 0;
    return nbytes(W->b, W->x, W->aux, W->tmp, W->rres, W->phi0, W->dr, W->dz, W->dp, W->dphi, W->dap,
                  W->omega, W->scale, W->st, W->part, W->flags) +
           work_bytes<NS, double>(W->w64.get()) + work_bytes<NS, float>(W->w32.get());
}

size_t operator_bytes(const m3d_operator& op) {
    size_t tiles = 0;
    for (const auto& T : op.tiles) tiles += nbytes(T.stage_off, T.stage_rows, T.slice_off, T.slc, T.slot, T.words);
    return tiles + nbytes(op.d_perm, op.l_col, op.u_col, op.hang, op.slc, op.l_val, op.u_val, op.diag_s, op.me,
                  op.l_val32, op.u_val32, op.diag_s32, op.me32, op.l_pk, op.u_pk, op.vtab, op.vtab32, op.bc,
                  op.constrained, op.nt_ptr, op.nt_edge, op.g_col, op.l_ptr, op.l_colidx, op.nt_val,
                  op.g_val, op.ms, op.l_vals, op.l_diag, op.l_vals32, op.l_diag32, op.dc_bc_edge,
                  op.dc_dead, op.dc_bc_node, op.m_adj_ptr, op.m_adj_cell, op.l_edge, op.d_err, op.m_area, op.m_len, op.m_massless,
                  op.m_dead, op.dc_in_pattern, op.s_cell_ptr, op.s_cell_edges, op.s_edge_ptr,
                  op.s_edge_cells, op.s_weight, op.s_sigma, op.s_omega, op.order[0], op.order[1],
                  op.order[2], op.order[3], op.order[4], op.order[5]);
}

size_t workspace_bytes(const m3d_operator& op) {
    return work64_bytes<1>(op.w1.get()) + work64_bytes<2>(op.w2.get()) + work64_bytes<4>(op.w4.get()) +
           work64_bytes<8>(op.w8.get()) + work64_bytes<16>(op.w16.get()) + work64_bytes<32>(op.w32.get());
}

void clear_graphs(m3d_operator& op) {
    if (op.w1) op.w1->graphs.clear();
    if (op.w2) op.w2->graphs.clear();
    if (op.w4) op.w4->graphs.clear();
    if (op.w8) op.w8->graphs.clear();
    if (op.w16) op.w16->graphs.clear();
    if (op.w32) op.w32->graphs.clear();
}

void set_mass(m3d_operator& op, const double* me) {
    clear_graphs(op);
    std::vector<double> m(op.npad, 0.0);
    for (int r = 0; r < op.npad; ++r) {
        const int e = op.perm[r];
        if (e >= 0 && !op.h_constrained[e]) m[r] = me[e];
    }
    op.me.upload(m, op.ctx->stream);
    std::vector<float> m32(m.begin(), m.end());
    op.me32.upload(m32, op.ctx->stream);
    HIPCHK(hipStreamSynchronize(op.ctx->stream));
    op.mass_set = true;
}

void set_divcorr(m3d_operator& op, const m3d_divcorr_desc& d) {
    clear_graphs(op);
    const int n = op.n, nn = d.nn;
    hipStream_t s = op.ctx->stream;
    if (nn <= 0) throw Fail(M3D_ERR_ARG, "DivCorr needs at least one node");
    // Node -> incident edges in ascending caller edge order.
    std::vector<int> ptr(nn + 1, 0);
    for (int e = 0; e < n; ++e)
        for (int k = 2 * e; k < 2 * e + 2; ++k) {
            const int v = d.g_col[k];
            if (v < 0 || v >= nn) throw Fail(M3D_ERR_ARG, "G column out of range");
            ptr[v + 1]++;
        }
    for (int v = 0; v < nn; ++v) ptr[v + 1] += ptr[v];
    std::vector<int> edge(ptr[nn]);
    std::vector<double> val(ptr[nn]);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < n; ++e)
        for (int k = 2 * e; k < 2 * e + 2; ++k) {
            const int v = d.g_col[k];
            const int q = pos[v]++;
            edge[q] = op.iperm[e];
            val[q] = d.g_val[k];
        }
    std::vector<int> gcol(2 * size_t(op.npad), -1);
    std::vector<double> gval(2 * size_t(op.npad), 0.0), ms(op.npad, 0.0);
    std::vector<unsigned char> bce(op.npad, 1), dead(op.npad, 1);
    for (int r = 0; r < op.npad; ++r) {
        const int e = op.perm[r];
        if (e < 0) continue;
        gcol[2 * r] = d.g_col[2 * e]; gcol[2 * r + 1] = d.g_col[2 * e + 1];
        gval[2 * r] = d.g_val[2 * e]; gval[2 * r + 1] = d.g_val[2 * e + 1];
        ms[r] = d.dead_edge[e] ? 0.0 : d.me_sigma[e];
        bce[r] = d.bc_edge[e];
        dead[r] = d.dead_edge[e];
    }
    op.nn = nn;
    op.nt_ptr.upload(ptr, s); op.nt_edge.upload(edge, s); op.nt_val.upload(val, s);
    op.g_col.upload(gcol, s); op.g_val.upload(gval, s); op.ms.upload(ms, s);
    op.dc_bc_edge.upload(bce, s); op.dc_dead.upload(dead, s);
    op.dc_bc_node.upload(std::vector<unsigned char>(d.bc_node, d.bc_node + nn), s);
    const int lnnz = d.L_rowptr[nn];
    op.l_ptr.upload(std::vector<int>(d.L_rowptr, d.L_rowptr + nn + 1), s);
    op.l_colidx.upload(std::vector<int>(d.L_colidx, d.L_colidx + lnnz), s);
    op.l_vals.upload(std::vector<double>(d.L_values, d.L_values + lnnz), s);
    op.l_diag.upload(std::vector<double>(d.L_diag, d.L_diag + nn), s);
    op.l32_fresh = false;
    // For sigma updates on the device: the edges that make up the pattern of
    // L (weight >= 1e-30, as div_laplacian keeps them) and, per off-diagonal
    // entry, the one edge joining its two nodes.  A pattern that is not of
    // that form (an edge with both ends on one node, two edges on one node
    // pair, an entry without an edge) leaves device updates unavailable.
    std::vector<unsigned char> pat(op.npad, 0);
    std::vector<int> ledge(std::max(lnnz, 1), -1);
    bool ok = true;
    auto find = [&](int row, int col) {
        for (int k = d.L_rowptr[row]; k < d.L_rowptr[row + 1]; ++k)
            if (d.L_colidx[k] == col) return k;
        return -1;
    };
    for (int e = 0; e < n && ok; ++e) {
        const double w = d.dead_edge[e] ? 0.0 : d.me_sigma[e];
        if (!(w >= 1e-30)) continue;
        pat[op.iperm[e]] = 1;
        const int a = d.g_col[2 * e], b = d.g_col[2 * e + 1];
        const int kab = a == b ? -1 : find(a, b), kba = a == b ? -1 : find(b, a);
        if (kab < 0 || kba < 0 || ledge[kab] >= 0 || ledge[kba] >= 0) {
            ok = false;
            break;
        }
        ledge[kab] = op.iperm[e];
        ledge[kba] = op.iperm[e];
    }
    for (int v = 0; v < nn && ok; ++v)
        for (int k = d.L_rowptr[v]; k < d.L_rowptr[v + 1]; ++k)
            if (d.L_colidx[k] != v && ledge[k] < 0) ok = false;
    op.dc_in_pattern.upload(pat, s);
    op.l_edge.upload(ledge, s);
    op.sigma_update_ok = ok;
    HIPCHK(hipStreamSynchronize(s));
    op.dc_set = true;
}

void set_mass_structure(m3d_operator& op, const m3d_mass_structure& d) {
    const int n = op.n;
    if (d.nc <= 0 || !d.adj_ptr || !d.dual_area || !d.length || !d.massless || !d.dead)
        throw Fail(M3D_ERR_ARG, "incomplete mass structure");
    if (d.adj_ptr[0] != 0) throw Fail(M3D_ERR_ARG, "adj_ptr must start at 0");
    for (int e = 0; e < n; ++e)
        if (d.adj_ptr[e + 1] < d.adj_ptr[e]) throw Fail(M3D_ERR_ARG, "adj_ptr must not decrease");
    const int na = d.adj_ptr[n];
    if (na > 0 && !d.adj_cell) throw Fail(M3D_ERR_ARG, "adj_cell missing");
    hipStream_t s = op.ctx->stream;
    op.mass_nc = d.nc;
    op.m_adj_ptr.upload(std::vector<int>(d.adj_ptr, d.adj_ptr + n + 1), s);
    op.m_adj_cell.upload(na > 0 ? std::vector<int>(d.adj_cell, d.adj_cell + na) : std::vector<int>(1, -1), s);
    op.m_area.upload(std::vector<double>(d.dual_area, d.dual_area + n), s);
    op.m_len.upload(std::vector<double>(d.length, d.length + n), s);
    op.m_massless.upload(std::vector<unsigned char>(d.massless, d.massless + n), s);
    op.m_dead.upload(std::vector<unsigned char>(d.dead, d.dead + n), s);
    if (!op.d_err.p) op.d_err.alloc(1);
    HIPCHK(hipStreamSynchronize(s));
    op.mass_struct = true;
}

// Me, the DivCorr weights and Laplacian values and the sensitivity sigma from
// the device array d_sigma [mass_nc], on the context stream.
void update_sigma(m3d_operator& op, const double* d_sigma) {
    if (!op.mass_struct) throw Fail(M3D_ERR_STATE, "m3d_operator_set_mass_structure has not been called");
    if (!op.dc_set) throw Fail(M3D_ERR_STATE, "m3d_operator_set_divcorr has not been called");
    if (!op.sigma_update_ok)
        throw Fail(M3D_ERR_STATE, "the DivCorr Laplacian pattern is not one edge per node pair; use the host update");
    hipStream_t s = op.ctx->stream;
    if (!op.me.p || op.me.n != size_t(op.npad)) op.me.alloc(op.npad);
    if (!op.me32.p || op.me32.n != size_t(op.npad)) op.me32.alloc(op.npad);
    HIPCHK(hipMemsetAsync(op.d_err.p, 0, sizeof(int), s));
    launch_sigma_edges(s, op.npad, op.mass_nc, op.d_perm.p, op.m_adj_ptr.p, op.m_adj_cell.p, op.m_area.p,
                       op.m_len.p, op.m_massless.p, op.m_dead.p, op.constrained.p, op.dc_dead.p,
                       op.dc_in_pattern.p, d_sigma, op.me.p, op.me32.p, op.ms.p, op.d_err.p);
    launch_sigma_laplacian(s, op.nn, op.nt_ptr.p, op.nt_edge.p, op.nt_val.p, op.ms.p, op.l_ptr.p, op.l_colidx.p,
                           op.l_edge.p, op.g_col.p, op.g_val.p, op.dc_bc_node.p, op.l_vals.p, op.l_diag.p);
    HIPCHK(hipGetLastError());
    op.l32_fresh = false;
    if (op.nc == op.mass_nc && op.s_sigma.p)
        HIPCHK(hipMemcpyAsync(op.s_sigma.p, d_sigma, size_t(op.nc) * sizeof(double), hipMemcpyDeviceToDevice, s));
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, op.d_err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    op.mass_set = true;
    if (err)
        throw Fail(M3D_ERR_STATE, "the new sigma changes the DivCorr Laplacian pattern (an edge weight crosses "
                                  "1e-30); refresh with m3d_operator_set_mass and m3d_operator_set_divcorr");
}

// Host staging for the host-memory entry points.
struct Staging {
    DBuf<double2> rhs, aux, x;
};

}  // namespace

extern "C" {

const char* m3d_last_error(void) { return last_error().c_str(); }

int m3d_context_create(int device, m3d_context** out) {
    return guarded([&] {
        if (!out) throw Fail(M3D_ERR_ARG, "null output pointer");
        int count = 0;
        HIPCHK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Fail(M3D_ERR_ARG, "device index out of range");
        HIPCHK(hipSetDevice(device));
        (void)device_scratch(device);  // the no-allocation scratch's address on this device
        auto* c = new m3d_context;
        c->device = device;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            throw Fail(M3D_ERR_HIP, "hipStreamCreate failed");
        }
        *out = c;
    });
}

int m3d_context_destroy(m3d_context* ctx) {
    return guarded([&] {
        if (!ctx || ctx->closed) return;
        context_close(ctx);
    });
}

void* m3d_context_stream(m3d_context* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int m3d_operator_create(m3d_context* ctx, const m3d_operator_desc* desc, m3d_operator** out) {
    return guarded([&] {
        if (!ctx || !desc || !out) throw Fail(M3D_ERR_ARG, "null argument");
        if (desc->n <= 0 || !desc->rowptr || !desc->colidx || !desc->values || !desc->diag_s ||
            !desc->constrained || !desc->bc)
            throw Fail(M3D_ERR_ARG, "incomplete operator description");
        if (desc->n_hanging > 0 && !desc->hanging) throw Fail(M3D_ERR_ARG, "hanging triples missing");
        const int nnz = desc->rowptr[desc->n];
        for (int k = 0; k < nnz; ++k)
            if (desc->colidx[k] < 0 || desc->colidx[k] >= desc->n) throw Fail(M3D_ERR_ARG, "column index out of range");
        HIPCHK(hipSetDevice(ctx->device));
        std::unique_ptr<m3d_operator> op(new m3d_operator);
        if (ctx->closed) throw Fail(M3D_ERR_STATE, "context already destroyed");
        op->ctx = ctx;
        build_operator(*op, *desc);
        *out = op.release();
        context_attach(ctx);
    });
}

int m3d_operator_destroy(m3d_operator* op) {
    return guarded([&] {
        if (!op) return;
        m3d_context* ctx = op->ctx;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        delete op;
        context_detach(ctx);
    });
}

int m3d_operator_set_context(m3d_operator* op, m3d_context* ctx) {
    return guarded([&] {
        if (!op || !ctx) throw Fail(M3D_ERR_ARG, "null argument");
        if (ctx->closed) throw Fail(M3D_ERR_STATE, "context already destroyed");
        m3d_context* old = op->ctx;
        if (ctx == old) return;
        if (ctx->device != old->device) throw Fail(M3D_ERR_ARG, "the context is on another device");
        HIPCHK(hipSetDevice(old->device));
        HIPCHK(hipStreamSynchronize(old->stream));
        context_attach(ctx);
        op->ctx = ctx;
        context_detach(old);
    });
}

int m3d_operator_set_mass(m3d_operator* op, const double* me) {
    return guarded([&] {
        if (!op || !me) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        set_mass(*op, me);
    });
}

int m3d_operator_set_divcorr(m3d_operator* op, const m3d_divcorr_desc* desc) {
    return guarded([&] {
        if (!op || !desc || !desc->g_col || !desc->g_val || !desc->me_sigma || !desc->bc_edge ||
            !desc->dead_edge || !desc->bc_node || !desc->L_rowptr || !desc->L_colidx ||
            !desc->L_values || !desc->L_diag)
            throw Fail(M3D_ERR_ARG, "incomplete DivCorr description");
        const int lnnz = desc->L_rowptr[desc->nn];
        for (int k = 0; k < lnnz; ++k)
            if (desc->L_colidx[k] < 0 || desc->L_colidx[k] >= desc->nn) throw Fail(M3D_ERR_ARG, "L column out of range");
        HIPCHK(hipSetDevice(op->ctx->device));
        set_divcorr(*op, *desc);
    });
}

int m3d_operator_set_mass_structure(m3d_operator* op, const m3d_mass_structure* desc) {
    return guarded([&] {
        if (!op || !desc) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        set_mass_structure(*op, *desc);
    });
}

int m3d_operator_update_sigma(m3d_operator* op, const double* sigma) {
    return guarded([&] {
        if (!op || !sigma) throw Fail(M3D_ERR_ARG, "null argument");
        if (!op->mass_struct) throw Fail(M3D_ERR_STATE, "m3d_operator_set_mass_structure has not been called");
        HIPCHK(hipSetDevice(op->ctx->device));
        DBuf<double> d;
        d.upload(std::vector<double>(sigma, sigma + op->mass_nc), op->ctx->stream);
        update_sigma(*op, d.p);
    });
}

int m3d_operator_update_sigma_device(m3d_operator* op, const double* d_sigma) {
    return guarded([&] {
        if (!op || !d_sigma) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        update_sigma(*op, d_sigma);
    });
}

int m3d_operator_set_sensitivity(m3d_operator* op, const m3d_sensitivity_desc* desc) {
    return guarded([&] {
        if (!op || !desc || desc->nc <= 0 || !desc->cell_ptr || !desc->cell_edges || !desc->edge_weight ||
            !desc->sigma)
            throw Fail(M3D_ERR_ARG, "incomplete sensitivity description");
        const int nc = desc->nc;
        const int total = desc->cell_ptr[nc];
        if (desc->cell_ptr[0] != 0 || total < 0) throw Fail(M3D_ERR_ARG, "cell_ptr must start at 0");
        for (int c = 0; c < nc; ++c)
            if (desc->cell_ptr[c + 1] < desc->cell_ptr[c]) throw Fail(M3D_ERR_ARG, "cell_ptr must not decrease");
        for (int k = 0; k < total; ++k)
            if (desc->cell_edges[k] < 0 || desc->cell_edges[k] >= op->n)
                throw Fail(M3D_ERR_ARG, "cell edge out of range");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        // Host copies stay alive until the stream has drained.
        const std::vector<int> ptr(desc->cell_ptr, desc->cell_ptr + nc + 1);
        const std::vector<int> edges(desc->cell_edges, desc->cell_edges + total);
        const std::vector<double> w(desc->edge_weight, desc->edge_weight + op->n);
        const std::vector<double> sg(desc->sigma, desc->sigma + nc);
        op->s_cell_ptr.upload(ptr, s);
        op->s_cell_edges.upload(edges, s);
        // The transpose: per edge, the cells whose edge lists hold it, in
        // ascending cell order (with multiplicity).
        std::vector<int> eptr(op->n + 1, 0);
        for (int k = 0; k < total; ++k) eptr[edges[k] + 1]++;
        for (int e = 0; e < op->n; ++e) eptr[e + 1] += eptr[e];
        std::vector<int> ecells(std::max(total, 1), 0), epos(eptr.begin(), eptr.end() - 1);
        for (int c = 0; c < nc; ++c)
            for (int k = ptr[c]; k < ptr[c + 1]; ++k) ecells[epos[edges[k]]++] = c;
        op->s_edge_ptr.upload(eptr, s);
        op->s_edge_cells.upload(ecells, s);
        op->s_weight.upload(w, s);
        op->s_sigma.upload(sg, s);
        HIPCHK(hipStreamSynchronize(s));
        op->nc = nc;
    });
}

int m3d_perturbation_rhs_device(m3d_operator* op, int nsys, const double* omega, const double* d_ebg,
                                const double* d_dsigma, double* d_rhs) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (op->nc <= 0) throw Fail(M3D_ERR_STATE, "m3d_operator_set_sensitivity has not been called");
        if (!op->mass_struct) throw Fail(M3D_ERR_STATE, "m3d_operator_set_mass_structure has not been called");
        if (op->mass_nc != op->nc) throw Fail(M3D_ERR_STATE, "mass structure and sensitivity disagree on cells");
        if (nsys <= 0 || !omega || !d_ebg || !d_dsigma || !d_rhs) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        const std::vector<double> om(omega, omega + nsys);
        op->s_omega.upload(om, s);
        launch_perturbation_rhs(s, op->n, nsys, op->nc, op->s_edge_ptr.p, op->s_edge_cells.p, op->m_adj_ptr.p,
                                op->m_adj_cell.p, op->m_len.p, op->s_sigma.p, d_dsigma, op->s_omega.p,
                                reinterpret_cast<const double2*>(d_ebg), reinterpret_cast<double2*>(d_rhs));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    });
}

int m3d_sensitivity_device(m3d_operator* op, int nsys, const double* omega, const double* d_E,
                           const double* d_lambda, double* d_grad) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (op->nc <= 0) throw Fail(M3D_ERR_STATE, "m3d_operator_set_sensitivity has not been called");
        if (nsys <= 0 || !omega || !d_E || !d_lambda || !d_grad) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        const std::vector<double> om(omega, omega + nsys);
        op->s_omega.upload(om, s);
        launch_sensitivity(s, op->nc, op->n, nsys, op->s_cell_ptr.p, op->s_cell_edges.p, op->s_weight.p,
                           op->s_sigma.p, op->s_omega.p, reinterpret_cast<const double2*>(d_E),
                           reinterpret_cast<const double2*>(d_lambda), d_grad);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    });
}

int m3d_sensitivity_device_noalloc(m3d_operator* op, int nsys, const double* omega, const double* d_E,
                                   const double* d_lambda, double* d_grad) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (op->nc <= 0) throw Fail(M3D_ERR_STATE, "m3d_operator_set_sensitivity has not been called");
        if (nsys <= 0 || !omega || !d_E || !d_lambda || !d_grad) throw Fail(M3D_ERR_ARG, "null argument");
        if (nsys > kNoAllocMaxSystems)
            throw Fail(M3D_ERR_ARG, "m3d_sensitivity_device_noalloc: " + std::to_string(nsys) +
                                        " systems exceed its scratch of " + std::to_string(kNoAllocMaxSystems));
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        double* d_omega = device_scratch(op->ctx->device);
        HIPCHK(hipMemcpyAsync(d_omega, omega, nsys * sizeof(double), hipMemcpyHostToDevice, s));
        launch_sensitivity(s, op->nc, op->n, nsys, op->s_cell_ptr.p, op->s_cell_edges.p, op->s_weight.p,
                           op->s_sigma.p, d_omega, reinterpret_cast<const double2*>(d_E),
                           reinterpret_cast<const double2*>(d_lambda), d_grad);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    });
}

int m3d_operator_permutation(m3d_operator* op, int* perm, int* colour_start) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (perm) std::copy(op->perm.begin(), op->perm.end(), perm);
        if (colour_start) std::copy(op->colour_start.begin(), op->colour_start.end(), colour_start);
    });
}

int m3d_operator_device_bytes(m3d_operator* op, long long* operator_bytes_out, long long* workspace_bytes_out) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (operator_bytes_out) *operator_bytes_out = static_cast<long long>(operator_bytes(*op));
        if (workspace_bytes_out) *workspace_bytes_out = static_cast<long long>(workspace_bytes(*op));
    });
}

int m3d_operator_trim(m3d_operator* op) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        HIPCHK(hipSetDevice(op->ctx->device));
        HIPCHK(hipStreamSynchronize(op->ctx->stream));
        op->w1.reset();
        op->w2.reset();
        op->w4.reset();
        op->w8.reset();
        op->w16.reset();
        op->w32.reset();
    });
}

int m3d_operator_format(m3d_operator* op, int* value_table, int* column_bits) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (value_table) *value_table = op->ntab;
        if (column_bits) *column_bits = op->ntab ? op->cbits : 0;
    });
}

int m3d_operator_info(m3d_operator* op, int* n_padded, int* ncolours, long long* stored_entries,
                      long long* nnz) {
    return guarded([&] {
        if (!op) throw Fail(M3D_ERR_ARG, "null operator");
        if (n_padded) *n_padded = op->npad;
        if (ncolours) *ncolours = op->ncolours;
        if (stored_entries) *stored_entries = op->stored;
        if (nnz) *nnz = op->nnz;
    });
}

int m3d_solve_batch_device(m3d_operator* op, int nsys, int mode, const double* omega,
                           const double* d_rhs, const double* d_aux, double* d_x,
                           const m3d_solve_params* params, m3d_solve_result* results) {
    return guarded([&] {
        check_ready(op);
        check_params(params);
        if (nsys <= 0 || !omega || !d_rhs || !d_x || !results) throw Fail(M3D_ERR_ARG, "null argument");
        if (mode == M3D_MODE_FORWARD && !d_aux) throw Fail(M3D_ERR_ARG, "FORWARD needs E_pri in aux");
        HIPCHK(hipSetDevice(op->ctx->device));
        solve_dispatch(op, nsys, mode, omega, reinterpret_cast<const double2*>(d_rhs),
                       reinterpret_cast<const double2*>(d_aux), reinterpret_cast<double2*>(d_x),
                       *params, results);
    });
}

int m3d_solve_batch(m3d_operator* op, int nsys, int mode, const double* omega, const double* rhs,
                    const double* aux, double* x, const m3d_solve_params* params,
                    m3d_solve_result* results) {
    return guarded([&] {
        check_ready(op);
        check_params(params);
        if (nsys <= 0 || !omega || !rhs || !x || !results) throw Fail(M3D_ERR_ARG, "null argument");
        if (mode == M3D_MODE_FORWARD && !aux) throw Fail(M3D_ERR_ARG, "FORWARD needs E_pri in aux");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        const size_t bytes = size_t(nsys) * op->n * sizeof(double2);
        Staging st;
        st.rhs.alloc(size_t(nsys) * op->n);
        st.x.alloc(size_t(nsys) * op->n);
        HIPCHK(hipMemcpyAsync(st.rhs.p, rhs, bytes, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(st.x.p, x, bytes, hipMemcpyHostToDevice, s));
        if (aux) {
            st.aux.alloc(size_t(nsys) * op->n);
            HIPCHK(hipMemcpyAsync(st.aux.p, aux, bytes, hipMemcpyHostToDevice, s));
        }
        solve_dispatch(op, nsys, mode, omega, st.rhs.p, aux ? st.aux.p : nullptr, st.x.p, *params, results);
        HIPCHK(hipMemcpyAsync(x, st.x.p, bytes, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

}  // extern "C"

namespace {

template <int NS>
void spmv_impl(m3d_operator& op, int nsys, const double* omega, const double2* d_x, double2* d_y) {
    Work64<NS>& W = ready_workspace<NS>(op);
    hipStream_t s = op.ctx->stream;
    std::vector<double> om(NS, 0.0);
    for (int i = 0; i < nsys; ++i) om[i] = omega[i];
    W.h_omega = om;
    HIPCHK(hipMemcpyAsync(W.omega.p, om.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
    Launch<NS, double>::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_x, W.x.p);
    Launch<NS, double>::spmv(s, op.view<double>(), W.omega.p, W.x.p, nullptr, W.tmp.p);
    Launch<NS, double>::unpack(s, op.npad, op.n, nsys, op.d_perm.p, W.tmp.p, d_y);
    HIPCHK(hipGetLastError());
}

template <int NS>
void ssor_impl(m3d_operator& op, int nsys, const double* omega, const double2* d_r, double2* d_z) {
    Work64<NS>& W = ready_workspace<NS>(op);
    auto& w = vec64<NS>(op, W);
    hipStream_t s = op.ctx->stream;
    std::vector<double> om(NS, 0.0);
    for (int i = 0; i < nsys; ++i) om[i] = omega[i];
    W.h_omega = om;
    HIPCHK(hipMemcpyAsync(W.omega.p, om.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
    const unsigned mask = (nsys >= 32) ? 0xFFFFFFFFu : ((1u << nsys) - 1u);
    launch_bicg_begin(s, W.st.p, mask, 1, 1.0);
    Launch<NS, double>::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_r, W.x.p);
    const DevOp<double> A = op.view<double>();
    ssor_sweeps<NS, double>(s, op, A, W.omega.p, kInPlain, W.x.p, nullptr, nullptr, w.y.p, W.tmp.p, W.st.p);
    Launch<NS, double>::unpack(s, op.npad, op.n, nsys, op.d_perm.p, W.tmp.p, d_z);
    launch_bicg_begin(s, W.st.p, 0u, 1, 1.0);
    HIPCHK(hipGetLastError());
}

template <int NS>
void divcorr_host_impl(m3d_operator& op, int nsys, const double* omega, double2* d_E,
                       const double2* d_phi0_sm, const m3d_solve_params& prm) {
    Work64<NS>& W = ready_workspace<NS>(op);
    ensure_dc_vectors<NS>(op, W);
    hipStream_t s = op.ctx->stream;
    std::vector<double> om(NS, 0.0);
    for (int i = 0; i < nsys; ++i) om[i] = omega[i];
    W.h_omega = om;
    HIPCHK(hipMemcpyAsync(W.omega.p, om.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
    Launch<NS, double>::pack(s, op.npad, op.n, nsys, op.d_perm.p, d_E, W.x.p);
    const unsigned mask = (nsys >= 32) ? 0xFFFFFFFFu : ((1u << nsys) - 1u);
    divcorr<NS>(op, W, W.x.p, d_phi0_sm, mask, prm);
    Launch<NS, double>::unpack(s, op.npad, op.n, nsys, op.d_perm.p, W.x.p, d_E);
    HIPCHK(hipGetLastError());
}

// Average duration of one launch of a kernel class, timed with HIP events on
// the context stream, plus its algorithmic HBM bytes (free rows n, stored
// off-diagonal nonzeros m, k systems):
//   SpMV        m*e + n*2*8 + 2*k*n*16
//   SSOR apply  m*e + n*4*8 + 4*k*n*16   (lower half forward, upper half backward)
// with e = 4 bytes per entry when the width gathers from packed blocks,
// 8 + 4 otherwise.
template <int NS>
void bench_impl(m3d_operator& op, int kind, int nsys, const double* omega, int reps, double* avg_ms,
                double* bytes, double* stored_bytes) {
    Work64<NS>& W = ready_workspace<NS>(op);
    auto& w = vec64<NS>(op, W);
    hipStream_t s = op.ctx->stream;
    std::vector<double> om(NS, 0.0);
    for (int i = 0; i < nsys; ++i) om[i] = omega[i];
    W.h_omega = om;
    HIPCHK(hipMemcpyAsync(W.omega.p, om.data(), NS * sizeof(double), hipMemcpyHostToDevice, s));
    const unsigned mask = (nsys >= 32) ? 0xFFFFFFFFu : ((1u << nsys) - 1u);
    const size_t nv = size_t(op.npad) * NS;
    const DevOp<double> A = op.view<double>();
    if (kind >= 2) {
        // BiCGStab iterations on b = 1, x = 0 with tol 0, so no system stops.
        std::vector<double2> ones(nv, make_double2(1.0, 0.0));
        HIPCHK(hipMemcpyAsync(W.tmp.p, ones.data(), nv * sizeof(double2), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        launch_bicg_begin(s, W.st.p, mask, 1 << 30, 0.0);
        if (kind != 3) {
            Launch<NS, double>::fill_zero(s, nv, W.x.p);
            Launch<NS, double>::bicg_init(s, A, W.omega.p, W.tmp.p, W.x.p, w.r.p, w.r0.p, W.st.p, W.part.p,
                                          1);
        } else {
            auto& w32 = vec32<NS>(op, W);
            std::vector<float2> ones32(nv, make_float2(1.f, 0.f));
            HIPCHK(hipMemcpyAsync(w32.b.p, ones32.data(), nv * sizeof(float2), hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
            Launch<NS, float>::fill_zero(s, nv, w32.x.p);
            Launch<NS, float>::bicg_init(s, op.view<float>(), W.omega.p, w32.b.p, w32.x.p, w32.r.p, w32.r0.p,
                                         W.st.p, W.part.p, 1);
        }
    } else {
        launch_bicg_begin(s, W.st.p, mask, 1, 1.0);
    }
    Launch<NS, double>::fill_zero(s, nv, W.x.p);
    const int bwd_last = op.ncolours - (op.last_upper_empty ? 2 : 1);
    auto one = [&]() {
        if (kind == 0) {
            Launch<NS, double>::spmv(s, A, W.omega.p, W.x.p, nullptr, W.tmp.p);
        } else if (kind == 2) {
            bicg_iteration<NS, double>(op, W, w, W.x.p);
        } else if (kind == 3) {
            auto& w32 = vec32<NS>(op, W);
            bicg_iteration<NS, float>(op, W, w32, w32.x.p);
        } else if (kind == 4) {
            // The backward launches an SSOR application makes.
            for (int c = bwd_last; c >= 0; --c)
                Launch<NS, double>::sweep_bwd(s, A, W.omega.p, op.colour_start[c], op.colour_start[c + 1],
                                              W.x.p, W.tmp.p, W.st.p);
        } else {
            ssor_sweeps<NS, double>(s, op, A, W.omega.p, kInPlain, W.x.p, nullptr, nullptr, w.y.p, W.tmp.p,
                                    W.st.p);
        }
    };
    one();  // warm
    // Launched as the solver launches its work: from a graph (kGraphIters
    // BiCGStab iterations, or all reps of the other kinds, per graph launch)
    // unless graphs are switched off, so the event time per launch is close
    // to the kernels' own duration.
    int per = 1;
    hipGraphExec_t ex = nullptr;
    if (use_graphs()) {
        per = (kind == 2 || kind == 3) ? kGraphIters : reps;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < per; ++k) one();
        HIPCHK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) throw Fail(M3D_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        HIPCHK(hipGraphLaunch(ex, s));  // warm
    }
    const int units = (reps + per - 1) / per;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    for (int r = 0; r < units; ++r) {
        if (ex) HIPCHK(hipGraphLaunch(ex, s));
        else one();
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ex) (void)hipGraphExecDestroy(ex);
    launch_bicg_begin(s, W.st.p, 0u, 1, 1.0);
    HIPCHK(hipStreamSynchronize(s));
    // Kind 4 reports one launch: the mean over the colours of a sweep.
    *avg_ms = double(ms) / (units * per) / (kind == 4 ? bwd_last + 1 : 1);
    long long nfree = 0, nfree_bwd = 0;
    for (int e = 0; e < op.n; ++e) nfree += op.h_constrained[e] ? 0 : 1;
    for (int r = 0; r < op.colour_start[bwd_last + 1]; ++r)
        nfree_bwd += (op.perm[r] >= 0 && !op.h_constrained[op.perm[r]]) ? 1 : 0;
    const double m = double(op.nnz), n = double(nfree), k = double(nsys);
    // Algorithmic bytes: value + column of the m off-diagonal nonzeros
    // (w_v + 4 bytes each), the (S_ii, Me_i) diagonal pair, and each vector
    // touched once.  stored_bytes: the same with the bytes per coupling of
    // the format the launch reads, one 4-byte word in the packed blocks and
    // the row tiles.
    const LaunchShape& shp = op.shape[kind == 3 ? 1 : 0][width_index(NS)];
    const bool pk = op.ntab && (shp.packed || (kind == 0 && shp.tiled && op.tiles[width_index(NS)].ntiles));
    const double nb = double(nfree_bwd);
    auto count = [&](bool stored) {
        auto ent = [&](double wv) { return stored && pk ? 4.0 : wv + 4.0; };
        auto spmv_b = [&](double wv, double wc) { return m * ent(wv) + 2.0 * n * wv + 2.0 * k * n * wc; };
        auto ssor_b = [&](double wv, double wc) { return m * ent(wv) + 4.0 * n * wv + 4.0 * k * n * wc; };
        // One right-preconditioned BiCGStab iteration: 2 SpMV, 2 SSOR and
        // 17 vector passes of the AXPY / dot work.
        auto iter_b = [&](double wv, double wc) {
            return 2.0 * spmv_b(wv, wc) + 2.0 * ssor_b(wv, wc) + 17.0 * k * n * wc;
        };
        // Backward sweep: the upper half of the couplings, the diagonal
        // pair, y in and z out of the rows it visits, per launch (one colour
        // on average).
        auto bwd_b = [&](double wv, double wc) {
            return (0.5 * m * ent(wv) + 2.0 * nb * wv + 2.0 * k * nb * wc) / (bwd_last + 1);
        };
        return kind == 0   ? spmv_b(8, 16)
               : kind == 1 ? ssor_b(8, 16)
               : kind == 2 ? iter_b(8, 16)
               : kind == 3 ? iter_b(4, 8)
                           : bwd_b(8, 16);
    };
    *bytes = count(false);
    if (stored_bytes) *stored_bytes = count(true);
}

bool valid_shape(int width, const LaunchShape& sh) {
    const int sub = sh.sweep_sub;
    const bool sub_ok = (sub == 1 || sub == 2 || sub == 4 || sub == 8) && width * sub <= 32;
    const int ch = sh.sweep_chunk;
    return sub_ok && sh.sweep_cap >= kXcds && sh.sweep_cap <= 65536 && sh.sweep_cap % kXcds == 0 &&
           sh.row_cap >= kXcds && sh.row_cap <= kMaxBlocks && sh.row_cap % kXcds == 0 &&
           (ch == 1 || ch == 2 || ch == 4 || ch == 8 || ch == 16) && (sh.packed == 0 || sh.packed == 1) &&
           (sh.wave == 0 || (sh.wave == 1 && width >= 8)) && (sh.tiled == 0 || (sh.tiled == 1 && width >= 8));
}

int width_slot(int width) {
    for (int i = 0; i < kWidths; ++i)
        if (width == (1 << i)) return i;
    throw Fail(M3D_ERR_ARG, "width must be 1, 2, 4, 8, 16 or 32");
}

// ms per BiCGStab iteration at batch width 1 << slot in precision p.
double time_iteration(m3d_operator& op, int p, int slot, const double* omega, int reps) {
    const int kind = p == 0 ? 2 : 3, k = 1 << slot;
    double ms = 0.0, bytes = 0.0;
    switch (k) {
        case 1: bench_impl<1>(op, kind, k, omega, reps, &ms, &bytes, nullptr); break;
        case 2: bench_impl<2>(op, kind, k, omega, reps, &ms, &bytes, nullptr); break;
        case 4: bench_impl<4>(op, kind, k, omega, reps, &ms, &bytes, nullptr); break;
        case 8: bench_impl<8>(op, kind, k, omega, reps, &ms, &bytes, nullptr); break;
        case 16: bench_impl<16>(op, kind, k, omega, reps, &ms, &bytes, nullptr); break;
        default: bench_impl<32>(op, kind, k, omega, reps, &ms, &bytes, nullptr); break;
    }
    return ms;
}

// Coordinate descent over the launch shape of one (precision, width): each
// candidate is timed twice (the faster counts) and replaces the incumbent
// only when at least 5 % faster.  A shape changes rounding, and with it the
// iteration count of the ill-conditioned low-frequency systems by thousands
// either way, so a gain inside the timing noise (2-3 %) is not taken.
double tune_one(m3d_operator& op, int p, int slot, const double* omega, int reps) {
    const int width = 1 << slot;
    auto time_shape = [&](const LaunchShape& c) {
        op.shape[p][slot] = c;
        clear_graphs(op);
        const double a = time_iteration(op, p, slot, omega, reps);
        const double b = time_iteration(op, p, slot, omega, reps);
        return std::min(a, b);
    };
    LaunchShape best = op.shape[p][slot];
    double tb = time_shape(best);
    auto consider = [&](LaunchShape c) {
        if (!valid_shape(width, c)) return;
        const double t = time_shape(c);
        if (t < 0.95 * tb) {
            tb = t;
            best = c;
        }
    };
    const LaunchShape base = best;
    for (int sub : {1, 2, 4, 8}) {
        if (sub == base.sweep_sub) continue;
        LaunchShape c = best;
        c.sweep_sub = sub;
        consider(c);
    }
    const LaunchShape after_sub = best;
    for (int cap : {1024, 2048, 4096, 8192, 16384}) {
        if (cap == after_sub.sweep_cap) continue;
        LaunchShape c = best;
        c.sweep_cap = cap;
        consider(c);
    }
    const LaunchShape after_cap = best;
    for (int cap : {512, 1024, 2048}) {
        if (cap == after_cap.row_cap) continue;
        LaunchShape c = best;
        c.row_cap = cap;
        consider(c);
    }
    const LaunchShape after_row = best;
    for (int ch : {1, 2, 4, 8}) {
        if (ch == after_row.sweep_chunk) continue;
        LaunchShape c = best;
        c.sweep_chunk = ch;
        consider(c);
    }
    if (op.ntab) {
        LaunchShape c = best;
        c.packed = 1 - c.packed;
        consider(c);
    }
    if ((1 << slot) >= 8) {
        LaunchShape c = best;
        c.wave = 1 - c.wave;
        consider(c);
    }
    op.shape[p][slot] = best;
    clear_graphs(op);
    return tb;
}

template <typename F>
void for_width(int nsys, F&& f) {
    if (nsys <= 0 || nsys > kMaxSys) throw Fail(M3D_ERR_ARG, "nsys must be in 1..32");
    f(nsys);
}

}  // namespace

extern "C" {

int m3d_spmv(m3d_operator* op, int nsys, const double* omega, const double* x, double* y) {
    return guarded([&] {
        check_ready(op);
        if (!omega || !x || !y) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        for_width(nsys, [&](int k) {
            const size_t bytes = size_t(k) * op->n * sizeof(double2);
            Staging st;
            st.x.alloc(size_t(k) * op->n);
            st.rhs.alloc(size_t(k) * op->n);
            HIPCHK(hipMemcpyAsync(st.x.p, x, bytes, hipMemcpyHostToDevice, s));
            if (k == 1) spmv_impl<1>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k == 2) spmv_impl<2>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k <= 4) spmv_impl<4>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k <= 8) spmv_impl<8>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k <= 16) spmv_impl<16>(*op, k, omega, st.x.p, st.rhs.p);
            else spmv_impl<32>(*op, k, omega, st.x.p, st.rhs.p);
            HIPCHK(hipMemcpyAsync(y, st.rhs.p, bytes, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        });
    });
}

int m3d_ssor_apply(m3d_operator* op, int nsys, const double* omega, const double* r, double* z) {
    return guarded([&] {
        check_ready(op);
        if (!omega || !r || !z) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        for_width(nsys, [&](int k) {
            const size_t bytes = size_t(k) * op->n * sizeof(double2);
            Staging st;
            st.x.alloc(size_t(k) * op->n);
            st.rhs.alloc(size_t(k) * op->n);
            HIPCHK(hipMemcpyAsync(st.x.p, r, bytes, hipMemcpyHostToDevice, s));
            if (k == 1) ssor_impl<1>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k == 2) ssor_impl<2>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k <= 4) ssor_impl<4>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k <= 8) ssor_impl<8>(*op, k, omega, st.x.p, st.rhs.p);
            else if (k <= 16) ssor_impl<16>(*op, k, omega, st.x.p, st.rhs.p);
            else ssor_impl<32>(*op, k, omega, st.x.p, st.rhs.p);
            HIPCHK(hipMemcpyAsync(z, st.rhs.p, bytes, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        });
    });
}

int m3d_bench_kernel(m3d_operator* op, int kind, int nsys, const double* omega, int reps,
                     double* avg_ms, double* alg_bytes, double* stored_bytes) {
    return guarded([&] {
        check_ready(op);
        if (!omega || !avg_ms || !alg_bytes || reps <= 0 || kind < 0 || kind > 4)
            throw Fail(M3D_ERR_ARG, "invalid benchmark arguments");
        HIPCHK(hipSetDevice(op->ctx->device));
        for_width(nsys, [&](int k) {
            if (k == 1) bench_impl<1>(*op, kind, k, omega, reps, avg_ms, alg_bytes, stored_bytes);
            else if (k == 2) bench_impl<2>(*op, kind, k, omega, reps, avg_ms, alg_bytes, stored_bytes);
            else if (k <= 4) bench_impl<4>(*op, kind, k, omega, reps, avg_ms, alg_bytes, stored_bytes);
            else if (k <= 8) bench_impl<8>(*op, kind, k, omega, reps, avg_ms, alg_bytes, stored_bytes);
            else if (k <= 16) bench_impl<16>(*op, kind, k, omega, reps, avg_ms, alg_bytes, stored_bytes);
            else bench_impl<32>(*op, kind, k, omega, reps, avg_ms, alg_bytes, stored_bytes);
        });
    });
}

int m3d_operator_get_shape(m3d_operator* op, int precision, int width, m3d_launch_shape* out) {
    return guarded([&] {
        if (!op || !out) throw Fail(M3D_ERR_ARG, "null argument");
        if (precision != M3D_PREC_FP64 && precision != M3D_PREC_MIXED) throw Fail(M3D_ERR_ARG, "unknown precision");
        const LaunchShape& sh = op->shape[precision == M3D_PREC_MIXED ? 1 : 0][width_slot(width)];
        out->sweep_sub = sh.sweep_sub;
        out->sweep_cap = sh.sweep_cap;
        out->row_cap = sh.row_cap;
        out->sweep_chunk = sh.sweep_chunk;
        out->packed = sh.packed;
        out->wave = sh.wave;
        out->tiled = sh.tiled;
    });
}

int m3d_operator_set_shape(m3d_operator* op, int precision, int width, const m3d_launch_shape* shape) {
    return guarded([&] {
        if (!op || !shape) throw Fail(M3D_ERR_ARG, "null argument");
        if (precision != M3D_PREC_FP64 && precision != M3D_PREC_MIXED) throw Fail(M3D_ERR_ARG, "unknown precision");
        const int slot = width_slot(width);
        LaunchShape sh;
        sh.sweep_sub = shape->sweep_sub;
        sh.sweep_cap = shape->sweep_cap;
        sh.row_cap = shape->row_cap;
        sh.sweep_chunk = shape->sweep_chunk;
        sh.packed = shape->packed;
        sh.wave = shape->wave;
        sh.tiled = shape->tiled;
        if (!valid_shape(width, sh)) throw Fail(M3D_ERR_ARG, "invalid launch shape for this width");
        HIPCHK(hipSetDevice(op->ctx->device));
        HIPCHK(hipStreamSynchronize(op->ctx->stream));
        op->shape[precision == M3D_PREC_MIXED ? 1 : 0][slot] = sh;
        clear_graphs(*op);
    });
}

int m3d_operator_autotune(m3d_operator* op, int widths_mask, const double* omega, int reps,
                          double* best_ms) {
    return guarded([&] {
        check_ready(op);
        if (!omega || reps <= 0 || (widths_mask & ~((1 << kWidths) - 1)) != 0)
            throw Fail(M3D_ERR_ARG, "invalid autotune arguments");
        HIPCHK(hipSetDevice(op->ctx->device));
        if (best_ms)
            for (int i = 0; i < 2 * kWidths; ++i) best_ms[i] = 0.0;
        for (int p = 0; p < 2; ++p)
            for (int slot = 0; slot < kWidths; ++slot) {
                if (!(widths_mask & (1 << slot))) continue;
                const double t = tune_one(*op, p, slot, omega, reps);
                if (best_ms) best_ms[p * kWidths + slot] = t;
            }
    });
}

int m3d_divcorr(m3d_operator* op, int nsys, const double* omega, double* E, const double* phi0,
                const m3d_solve_params* params) {
    return guarded([&] {
        check_ready(op);
        check_params(params);
        if (!op->dc_set) throw Fail(M3D_ERR_STATE, "m3d_operator_set_divcorr has not been called");
        if (!omega || !E) throw Fail(M3D_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(op->ctx->device));
        hipStream_t s = op->ctx->stream;
        for_width(nsys, [&](int k) {
            int NS = 1;
            while (NS < k) NS *= 2;
            const size_t bytes = size_t(k) * op->n * sizeof(double2);
            Staging st;
            st.x.alloc(size_t(k) * op->n);
            HIPCHK(hipMemcpyAsync(st.x.p, E, bytes, hipMemcpyHostToDevice, s));
            // phi0 to the system-minor node layout.
            std::vector<double2> ph;
            if (phi0) {
                ph.assign(size_t(op->nn) * NS, make_double2(0.0, 0.0));
                for (int sy = 0; sy < k; ++sy)
                    for (int v = 0; v < op->nn; ++v)
                        ph[size_t(v) * NS + sy] = make_double2(phi0[2 * (size_t(sy) * op->nn + v)],
                                                               phi0[2 * (size_t(sy) * op->nn + v) + 1]);
                st.aux.alloc(ph.size());
                HIPCHK(hipMemcpyAsync(st.aux.p, ph.data(), ph.size() * sizeof(double2), hipMemcpyHostToDevice, s));
            }
            const double2* p0 = phi0 ? st.aux.p : nullptr;
            m3d_solve_params prm = *params;
            prm.divcorr_pcg_tol = 0.0;  // divcorr_poisson_ as the reference runs it
            if (NS == 1) divcorr_host_impl<1>(*op, k, omega, st.x.p, p0, prm);
            else if (NS == 2) divcorr_host_impl<2>(*op, k, omega, st.x.p, p0, prm);
            else if (NS == 4) divcorr_host_impl<4>(*op, k, omega, st.x.p, p0, prm);
            else if (NS == 8) divcorr_host_impl<8>(*op, k, omega, st.x.p, p0, prm);
            else if (NS == 16) divcorr_host_impl<16>(*op, k, omega, st.x.p, p0, prm);
            else divcorr_host_impl<32>(*op, k, omega, st.x.p, p0, prm);
            HIPCHK(hipMemcpyAsync(E, st.x.p, bytes, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        });
    });
}

}  // extern "C"
