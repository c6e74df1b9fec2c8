// Station side of the misfit gradient on the device: impedance extraction,
// weighted residual and misfit, and the adjoint right-hand side, on fields
// that stay in HBM (C-ABI in include/m3d_device.h, "Station side").
//
// The work is a few hundred stations with a dozen edges each, so every
// kernel is one thread per (frequency, station) or per (frequency, station
// edge); what matters is that the fields never leave the device between the
// forward and the adjoint batch.  Each formula follows the reference's
// operation order (forward_solver_fv.cpp:1534-1712, :1865-2008;
// inversion_fv.cpp:75-98, :258-271), with contraction into fused
// multiply-adds switched off, and the adjoint right-hand side is gathered per
// station edge over a precomputed contribution list in the reference's
// order of addition, so no two threads write one element and the sums are
// run-to-run identical.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "abi_common.h"
#include "dev_common.h"
#include "m3d_device.h"

using namespace m3d::abi;
using m3d::dev::cabs;
using m3d::dev::cdiv;
using m3d::dev::cmul;
using m3d::dev::mk;

namespace {

constexpr double kMu0 = 4.0e-7 * 3.14159265358979323846;
constexpr int kCache = 9;   // Hx1, Hy1, Hx2, Hy2, det H, Zxx, Zxy, Zyx, Zyy
constexpr int kCoef = 11;   // sx1, sy1, sx2, sy2, aHx1, aHy1, aHx2, aHy2, iz, ix, iy
constexpr int kThreads = 64;

__device__ inline double2 cadd(double2 a, double2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ inline double2 cscale(double2 a, double s) { return mk(a.x * s, a.y * s); }
__device__ inline double2 cdivr(double2 a, double s) { return mk(a.x / s, a.y / s); }

// Principal square root (the branch std::sqrt of a complex takes).
__device__ inline double2 csqrt_principal(double2 z) {
    if (z.x == 0.0 && z.y == 0.0) return mk(0.0, z.y);
    const double t = sqrt(2.0 * (cabs(z) + fabs(z.x)));
    if (z.x > 0.0) return mk(0.5 * t, z.y / t);
    return mk(fabs(z.y) / t, copysign(0.5 * t, z.y));
}

// extract_impedance_ per (frequency, station): group averages of the station
// cell's edges (:1534-1572), cell-centre E (:1577-1584), H from the
// finite-difference curl over i omega mu0 (:1597-1632), Z = E H^-1
// (:1699-1703) or the 1-D fallback (:1704-1712).
__global__ void k_station_z(int nf, int ns, int n, const int* __restrict__ valid, const int* __restrict__ eptr,
                            const int* __restrict__ edge, const unsigned char* __restrict__ role,
                            const double* __restrict__ csize, const double* __restrict__ ssig,
                            const double* __restrict__ omega, const double2* __restrict__ E,
                            double2* __restrict__ Z, double2* __restrict__ cache) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * ns) return;
    const int f = t / ns, s = t % ns;
    double2* z = Z + size_t(t) * 4;
    double2* hc = cache + size_t(t) * kCache;
    if (!valid[s]) {
        for (int k = 0; k < 4; ++k) z[k] = mk(0.0, 0.0);
        for (int k = 0; k < kCache; ++k) hc[k] = mk(0.0, 0.0);
        return;
    }
    const double2* e1 = E + size_t(2 * f) * n;
    const double2* e2 = E + size_t(2 * f + 1) * n;
    double2 x1t = mk(0, 0), x1b = mk(0, 0), x2t = mk(0, 0), x2b = mk(0, 0);
    double2 y1t = mk(0, 0), y1b = mk(0, 0), y2t = mk(0, 0), y2b = mk(0, 0);
    double2 z1px = mk(0, 0), z1mx = mk(0, 0), z2px = mk(0, 0), z2mx = mk(0, 0);
    double2 z1py = mk(0, 0), z1my = mk(0, 0), z2py = mk(0, 0), z2my = mk(0, 0);
    int nxt = 0, nxb = 0, nyt = 0, nyb = 0, nzpx = 0, nzmx = 0, nzpy = 0, nzmy = 0;
    for (int k = eptr[s]; k < eptr[s + 1]; ++k) {
        const int e = edge[k];
        const int r = role[k];
        const double2 a = e1[e], b = e2[e];
        if ((r & 3) == 0) {
            if (r & 4) { x1t = cadd(x1t, a); x2t = cadd(x2t, b); ++nxt; }
            else { x1b = cadd(x1b, a); x2b = cadd(x2b, b); ++nxb; }
        } else if ((r & 3) == 1) {
            if (r & 4) { y1t = cadd(y1t, a); y2t = cadd(y2t, b); ++nyt; }
            else { y1b = cadd(y1b, a); y2b = cadd(y2b, b); ++nyb; }
        } else {
            if (r & 4) { z1px = cadd(z1px, a); z2px = cadd(z2px, b); ++nzpx; }
            else { z1mx = cadd(z1mx, a); z2mx = cadd(z2mx, b); ++nzmx; }
            if (r & 8) { z1py = cadd(z1py, a); z2py = cadd(z2py, b); ++nzpy; }
            else { z1my = cadd(z1my, a); z2my = cadd(z2my, b); ++nzmy; }
        }
    }
    auto avg = [](double2& a, double2& b, int cnt) {
        if (cnt > 0) { a = cdivr(a, double(cnt)); b = cdivr(b, double(cnt)); }
    };
    avg(x1t, x2t, nxt); avg(x1b, x2b, nxb); avg(y1t, y2t, nyt); avg(y1b, y2b, nyb);
    avg(z1px, z2px, nzpx); avg(z1mx, z2mx, nzmx); avg(z1py, z2py, nzpy); avg(z1my, z2my, nzmy);
    auto centre = [](double2 top, double2 bot, int nt, int nb) {
        if (nt > 0 && nb > 0) return cscale(cadd(top, bot), 0.5);
        return nt > 0 ? top : bot;
    };
    const double2 Ex1 = centre(x1t, x1b, nxt, nxb), Ey1 = centre(y1t, y1b, nyt, nyb);
    const double2 Ex2 = centre(x2t, x2b, nxt, nxb), Ey2 = centre(y2t, y2b, nyt, nyb);
    const double cdx = csize[3 * s], cdy = csize[3 * s + 1], cdz = csize[3 * s + 2];
    double2 dEx1dz = mk(0, 0), dEx2dz = mk(0, 0), dEy1dz = mk(0, 0), dEy2dz = mk(0, 0);
    double2 dEz1dx = mk(0, 0), dEz2dx = mk(0, 0), dEz1dy = mk(0, 0), dEz2dy = mk(0, 0);
    if (nxt > 0 && nxb > 0) { dEx1dz = cdivr(csub(x1t, x1b), cdz); dEx2dz = cdivr(csub(x2t, x2b), cdz); }
    if (nyt > 0 && nyb > 0) { dEy1dz = cdivr(csub(y1t, y1b), cdz); dEy2dz = cdivr(csub(y2t, y2b), cdz); }
    if (nzpx > 0 && nzmx > 0) { dEz1dx = cdivr(csub(z1px, z1mx), cdx); dEz2dx = cdivr(csub(z2px, z2mx), cdx); }
    if (nzpy > 0 && nzmy > 0) { dEz1dy = cdivr(csub(z1py, z1my), cdy); dEz2dy = cdivr(csub(z2py, z2my), cdy); }
    const double2 iwmu = mk(0.0, omega[f] * kMu0);
    const double2 Hx1 = cdiv(csub(dEz1dy, dEy1dz), iwmu);
    const double2 Hy1 = cdiv(csub(dEx1dz, dEz1dx), iwmu);
    const double2 Hx2 = cdiv(csub(dEz2dy, dEy2dz), iwmu);
    const double2 Hy2 = cdiv(csub(dEx2dz, dEz2dx), iwmu);
    const double2 det = csub(cmul(Hx1, Hy2), cmul(Hx2, Hy1));
    double2 Zxx, Zxy, Zyx, Zyy;
    if (cabs(det) > 1e-30) {
        Zxx = cdiv(csub(cmul(Ex1, Hy2), cmul(Ex2, Hy1)), det);
        Zxy = cdiv(csub(cmul(Ex2, Hx1), cmul(Ex1, Hx2)), det);
        Zyx = cdiv(csub(cmul(Ey1, Hy2), cmul(Ey2, Hy1)), det);
        Zyy = cdiv(csub(cmul(Ey2, Hx1), cmul(Ey1, Hx2)), det);
    } else {
        const double2 z1d = csqrt_principal(cdiv(iwmu, mk(ssig[s], 0.0)));
        Zxx = mk(0.0, 0.0);
        Zxy = z1d;
        Zyx = mk(-z1d.x, -z1d.y);
        Zyy = mk(0.0, 0.0);
    }
    z[0] = Zxx; z[1] = Zxy; z[2] = Zyx; z[3] = Zyy;
    hc[0] = Hx1; hc[1] = Hy1; hc[2] = Hx2; hc[3] = Hy2; hc[4] = det;
    hc[5] = Zxx; hc[6] = Zxy; hc[7] = Zyx; hc[8] = Zyy;
}

// wr = w * w * conj(obs - pred) per datum (inversion_fv.cpp:258-271).
__global__ void k_station_wr(int count, const double2* __restrict__ obs, const double* __restrict__ err,
                             const double* __restrict__ wgt, const double2* __restrict__ cache, int ns,
                             double2* __restrict__ wr) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int fs = t / 4, c = t % 4;
    const double e = err[t];
    const double g = wgt ? wgt[t] : 1.0;
    if (g <= 0.0 || e <= 0.0) {
        wr[t] = mk(0.0, 0.0);
        return;
    }
    const double2 r = csub(obs[t], cache[size_t(fs) * kCache + 5 + c]);
    const double w = 1.0 / e;
    const double ww = w * w;
    wr[t] = mk(ww * r.x, ww * -r.y);
    (void)ns;
}

// phi = sum_f sum_s sum_c 0.5 w w |obs - pred|^2 in the reference's order
// (data_misfit_from_predicted_, inversion_fv.cpp:75-98): one thread, a few
// thousand terms.
__global__ void k_station_phi(int count, const double2* __restrict__ obs, const double* __restrict__ err,
                              const double* __restrict__ wgt, const double2* __restrict__ cache,
                              double* __restrict__ phi) {
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double acc = 0.0;
    for (int t = 0; t < count; ++t) {
        const double e = err[t];
        const double g = wgt ? wgt[t] : 1.0;
        if (g <= 0.0 || e <= 0.0) continue;
        const double2 r = csub(obs[t], cache[size_t(t / 4) * kCache + 5 + t % 4]);
        const double w = 1.0 / e;
        acc += 0.5 * w * w * (r.x * r.x + r.y * r.y);
    }
    phi[0] = acc;
}

// Station coefficients of build_adjoint_rhs_from_residual (:1865-1877): the
// E-part weights s = H^-T wr, the H-part weights aH = -Z^T s, and
// i_axis = 1 / (i omega mu0 h_axis); skip[t] = 1 where the reference skips
// the station (no cell, or |det H| < 1e-30).
__global__ void k_station_coef(int nf, int ns, const int* __restrict__ valid, const double* __restrict__ csize,
                               const double* __restrict__ omega, const double2* __restrict__ wr,
                               const double2* __restrict__ cache, double2* __restrict__ coef,
                               int* __restrict__ skip) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * ns) return;
    const int f = t / ns, s = t % ns;
    const double2* hc = cache + size_t(t) * kCache;
    const double2 det = hc[4];
    if (!valid[s] || cabs(det) < 1e-30) {
        skip[t] = 1;
        return;
    }
    skip[t] = 0;
    const double2 Hx1 = hc[0], Hy1 = hc[1], Hx2 = hc[2], Hy2 = hc[3];
    const double2 Zxx = hc[5], Zxy = hc[6], Zyx = hc[7], Zyy = hc[8];
    const double2* w = wr + size_t(t) * 4;
    const double2 sx1 = cdiv(csub(cmul(w[0], Hy2), cmul(w[1], Hx2)), det);
    const double2 sy1 = cdiv(csub(cmul(w[2], Hy2), cmul(w[3], Hx2)), det);
    const double2 sx2 = cdiv(cadd(cmul(mk(-w[0].x, -w[0].y), Hy1), cmul(w[1], Hx1)), det);
    const double2 sy2 = cdiv(cadd(cmul(mk(-w[2].x, -w[2].y), Hy1), cmul(w[3], Hx1)), det);
    const double2 aHx1 = cadd(cmul(Zxx, sx1), cmul(Zyx, sy1));
    const double2 aHy1 = cadd(cmul(Zxy, sx1), cmul(Zyy, sy1));
    const double2 aHx2 = cadd(cmul(Zxx, sx2), cmul(Zyx, sy2));
    const double2 aHy2 = cadd(cmul(Zxy, sx2), cmul(Zyy, sy2));
    const double wm = omega[f] * kMu0;
    double2* c = coef + size_t(t) * kCoef;
    c[0] = sx1; c[1] = sy1; c[2] = sx2; c[3] = sy2;
    c[4] = mk(-aHx1.x, -aHx1.y); c[5] = mk(-aHy1.x, -aHy1.y);
    c[6] = mk(-aHx2.x, -aHx2.y); c[7] = mk(-aHy2.x, -aHy2.y);
    // 1 / (i omega mu0 h) for h = dz, dx, dy.
    c[8] = cdiv(mk(1.0, 0.0), mk(0.0 * csize[3 * s + 2], wm * csize[3 * s + 2]));
    c[9] = cdiv(mk(1.0, 0.0), mk(0.0 * csize[3 * s], wm * csize[3 * s]));
    c[10] = cdiv(mk(1.0, 0.0), mk(0.0 * csize[3 * s + 1], wm * csize[3 * s + 1]));
}

__global__ void k_zero(size_t count, double2* __restrict__ x) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        x[i] = mk(0.0, 0.0);
}

// rhs of both polarisations on station edge k of frequency f: the station
// contributions of :1928-2008 added in the reference's order.
__global__ void k_station_rhs(int nf, int nu, int n, int ns, const int* __restrict__ uedge,
                              const int* __restrict__ cptr, const int4* __restrict__ cinfo,
                              const double* __restrict__ cwe, const double2* __restrict__ coef,
                              const int* __restrict__ skip, double2* __restrict__ rhs) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * nu) return;
    const int f = t / nu, k = t % nu;
    double2 v1 = mk(0.0, 0.0), v2 = mk(0.0, 0.0);
    for (int j = cptr[k]; j < cptr[k + 1]; ++j) {
        const int4 ci = cinfo[j];
        const int fs = f * ns + ci.x;
        if (skip[fs]) continue;
        const double2* c = coef + size_t(fs) * kCoef;
        if (ci.y == 0 || ci.y == 1) {
            const double we = cwe[j];
            const double2 a = ci.y == 0 ? c[0] : c[1];
            const double2 b = ci.y == 0 ? c[2] : c[3];
            v1 = cadd(v1, cscale(a, we));
            v2 = cadd(v2, cscale(b, we));
        } else {
            double2 ia = c[8 + (ci.z >> 1)];
            if (ci.z & 1) ia = mk(-ia.x, -ia.y);
            const double2 wh = cdivr(ia, double(ci.w));
            const double2 a = ci.y == 2 ? c[5] : c[4];
            const double2 b = ci.y == 2 ? c[7] : c[6];
            v1 = cadd(v1, cmul(a, wh));
            v2 = cadd(v2, cmul(b, wh));
        }
    }
    const int e = uedge[k];
    rhs[size_t(2 * f) * n + e] = v1;
    rhs[size_t(2 * f + 1) * n + e] = v2;
}

int blocks_for(long long count) { return int((count + kThreads - 1) / kThreads); }

}  // namespace

struct m3d_stations {
    m3d_context* ctx = nullptr;
    int n = 0, ns = 0, max_nf = 0, nu = 0, last_nf = 0;
    DBuf<int> valid, eptr, edge, uedge, cptr;
    DBuf<unsigned char> role;
    DBuf<double> csize, cwe, ssig, omega;
    DBuf<int4> cinfo;
    DBuf<double2> cache, coef;
    DBuf<int> skip;
    std::vector<double> h_omega;
};

extern "C" {

int m3d_stations_create(m3d_context* ctx, const m3d_station_desc* d, m3d_stations** out) {
    return guarded([&] {
        if (!ctx || !d || !out) throw Fail(M3D_ERR_ARG, "null argument");
        if (ctx->closed) throw Fail(M3D_ERR_STATE, "context already destroyed");
        if (d->n <= 0 || d->ns < 0 || d->max_nf <= 0 || d->nu < 0)
            throw Fail(M3D_ERR_ARG, "station table sizes out of range");
        if (d->ns > 0 && (!d->valid || !d->edge_ptr || !d->cell_size))
            throw Fail(M3D_ERR_ARG, "incomplete station table");
        if (d->nu > 0 && (!d->uedge || !d->contrib_ptr)) throw Fail(M3D_ERR_ARG, "incomplete station table");
        const int ne = d->ns > 0 ? d->edge_ptr[d->ns] : 0;
        const int nc = d->nu > 0 ? d->contrib_ptr[d->nu] : 0;
        if (ne > 0 && (!d->edge || !d->role)) throw Fail(M3D_ERR_ARG, "incomplete station table");
        if (nc > 0 && (!d->contrib || !d->contrib_we)) throw Fail(M3D_ERR_ARG, "incomplete station table");
        // Every index the kernels follow is checked here, so they read in bounds.
        for (int s = 0; s < d->ns; ++s)
            if (d->edge_ptr[s] > d->edge_ptr[s + 1]) throw Fail(M3D_ERR_ARG, "station edge pointers decrease");
        for (int k = 0; k < ne; ++k)
            if (d->edge[k] < 0 || d->edge[k] >= d->n) throw Fail(M3D_ERR_ARG, "station edge out of range");
        for (int k = 0; k < d->nu; ++k) {
            if (d->uedge[k] < 0 || d->uedge[k] >= d->n) throw Fail(M3D_ERR_ARG, "station edge out of range");
            if (d->contrib_ptr[k] > d->contrib_ptr[k + 1]) throw Fail(M3D_ERR_ARG, "contribution pointers decrease");
        }
        for (int j = 0; j < nc; ++j) {
            const int* c = d->contrib + 4 * j;
            if (c[0] < 0 || c[0] >= d->ns || c[1] < 0 || c[1] > 3 || c[2] < 0 || c[2] > 5 ||
                (c[1] >= 2 && c[3] <= 0))
                throw Fail(M3D_ERR_ARG, "malformed contribution");
        }
        HIPCHK(hipSetDevice(ctx->device));
        std::unique_ptr<m3d_stations> st(new m3d_stations);
        st->ctx = ctx;
        st->n = d->n;
        st->ns = d->ns;
        st->max_nf = d->max_nf;
        st->nu = d->nu;
        hipStream_t s = ctx->stream;
        st->valid.upload(std::vector<int>(d->valid, d->valid + d->ns), s);
        st->eptr.upload(d->ns > 0 ? std::vector<int>(d->edge_ptr, d->edge_ptr + d->ns + 1) : std::vector<int>{0}, s);
        st->edge.upload(std::vector<int>(d->edge, d->edge + ne), s);
        st->role.upload(std::vector<unsigned char>(d->role, d->role + ne), s);
        st->csize.upload(std::vector<double>(d->cell_size, d->cell_size + 3 * size_t(d->ns)), s);
        st->uedge.upload(std::vector<int>(d->uedge, d->uedge + d->nu), s);
        st->cptr.upload(d->nu > 0 ? std::vector<int>(d->contrib_ptr, d->contrib_ptr + d->nu + 1)
                                  : std::vector<int>{0}, s);
        std::vector<int4> ci(nc);
        for (int j = 0; j < nc; ++j)
            ci[j] = make_int4(d->contrib[4 * j], d->contrib[4 * j + 1], d->contrib[4 * j + 2], d->contrib[4 * j + 3]);
        st->cinfo.upload(ci, s);
        st->cwe.upload(std::vector<double>(d->contrib_we, d->contrib_we + nc), s);
        st->ssig.alloc(std::max(d->ns, 1));
        st->omega.alloc(d->max_nf);
        const size_t fs = size_t(d->max_nf) * std::max(d->ns, 1);
        st->cache.alloc(fs * kCache);
        st->coef.alloc(fs * kCoef);
        st->skip.alloc(fs);
        HIPCHK(hipStreamSynchronize(s));
        context_attach(ctx);
        *out = st.release();
    });
}

int m3d_stations_destroy(m3d_stations* st) {
    return guarded([&] {
        if (!st) return;
        m3d_context* ctx = st->ctx;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        delete st;
        context_detach(ctx);
    });
}

int m3d_stations_set_context(m3d_stations* st, m3d_context* ctx) {
    return guarded([&] {
        if (!st || !ctx) throw Fail(M3D_ERR_ARG, "null argument");
        if (ctx->closed) throw Fail(M3D_ERR_STATE, "context already destroyed");
        m3d_context* old = st->ctx;
        if (ctx == old) return;
        if (ctx->device != old->device) throw Fail(M3D_ERR_ARG, "the context is on another device");
        HIPCHK(hipSetDevice(old->device));
        HIPCHK(hipStreamSynchronize(old->stream));
        context_attach(ctx);
        st->ctx = ctx;
        context_detach(old);
    });
}

int m3d_impedance_device(m3d_stations* st, int nf, const double* omega, const double* station_sigma,
                         const double* d_E, double* d_Z) {
    return guarded([&] {
        if (!st || !omega || !d_E || !d_Z || (st->ns > 0 && !station_sigma)) throw Fail(M3D_ERR_ARG, "null argument");
        if (nf <= 0 || nf > st->max_nf) throw Fail(M3D_ERR_ARG, "frequency count outside 1..max_nf");
        HIPCHK(hipSetDevice(st->ctx->device));
        hipStream_t s = st->ctx->stream;
        st->h_omega.assign(omega, omega + nf);
        HIPCHK(hipMemcpyAsync(st->omega.p, omega, nf * sizeof(double), hipMemcpyHostToDevice, s));
        if (st->ns > 0) {
            HIPCHK(hipMemcpyAsync(st->ssig.p, station_sigma, st->ns * sizeof(double), hipMemcpyHostToDevice, s));
            k_station_z<<<blocks_for((long long)nf * st->ns), kThreads, 0, s>>>(
                nf, st->ns, st->n, st->valid.p, st->eptr.p, st->edge.p, st->role.p, st->csize.p, st->ssig.p,
                st->omega.p, reinterpret_cast<const double2*>(d_E), reinterpret_cast<double2*>(d_Z), st->cache.p);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(s));
        st->last_nf = nf;
    });
}

int m3d_weighted_residual_device(m3d_stations* st, int nf, const double* d_obs, const double* d_err,
                                 const double* d_weight, double* d_wr, double* d_phi) {
    return guarded([&] {
        if (!st || !d_obs || !d_err || !d_wr || !d_phi) throw Fail(M3D_ERR_ARG, "null argument");
        if (st->last_nf == 0) throw Fail(M3D_ERR_STATE, "m3d_impedance_device has not been called");
        if (nf != st->last_nf) throw Fail(M3D_ERR_ARG, "frequency count differs from the last impedance call");
        HIPCHK(hipSetDevice(st->ctx->device));
        hipStream_t s = st->ctx->stream;
        const int count = 4 * nf * st->ns;
        const auto* obs = reinterpret_cast<const double2*>(d_obs);
        if (count > 0)
            k_station_wr<<<blocks_for(count), kThreads, 0, s>>>(count, obs, d_err, d_weight, st->cache.p, st->ns,
                                                                 reinterpret_cast<double2*>(d_wr));
        k_station_phi<<<1, 1, 0, s>>>(count, obs, d_err, d_weight, st->cache.p, d_phi);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    });
}

int m3d_adjoint_rhs_device(m3d_stations* st, int nf, const double* d_wr, double* d_rhs) {
    return guarded([&] {
        if (!st || !d_wr || !d_rhs) throw Fail(M3D_ERR_ARG, "null argument");
        if (st->last_nf == 0) throw Fail(M3D_ERR_STATE, "m3d_impedance_device has not been called");
        if (nf != st->last_nf) throw Fail(M3D_ERR_ARG, "frequency count differs from the last impedance call");
        HIPCHK(hipSetDevice(st->ctx->device));
        hipStream_t s = st->ctx->stream;
        auto* rhs = reinterpret_cast<double2*>(d_rhs);
        const size_t total = size_t(2 * nf) * st->n;
        k_zero<<<int(std::min<size_t>((total + 255) / 256, 4096)), 256, 0, s>>>(total, rhs);
        if (st->ns > 0) {
            k_station_coef<<<blocks_for((long long)nf * st->ns), kThreads, 0, s>>>(
                nf, st->ns, st->valid.p, st->csize.p, st->omega.p, reinterpret_cast<const double2*>(d_wr),
                st->cache.p, st->coef.p, st->skip.p);
        }
        if (st->nu > 0) {
            k_station_rhs<<<blocks_for((long long)nf * st->nu), kThreads, 0, s>>>(
                nf, st->nu, st->n, st->ns, st->uedge.p, st->cptr.p, st->cinfo.p, st->cwe.p, st->coef.p, st->skip.p,
                rhs);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    });
}

}  // extern "C"
