// CPU oracle of the octree FV forward / adjoint solve.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithm on
// the CPU, in FP64, so tests can check the device solve against it; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product never calls it.
//
// It starts from the staggered grid (edges, faces, hanging lists, cell data)
// and re-derives everything after that independently of the product's host
// code, in the shape of the reference:
//   operators      src/octree/operators.cpp:19-170, 176-286, 344-466,
//                  481-551, 556-837 (the disabled grad-div block skipped),
//                  849-944, 952-1019, 1025-1074
//   CSR matvec     include/maple3dmt/common.h:52-61
//   SSOR           include/maple3dmt/forward/ssor.h:28-79 (natural order)
//   BiCGStab       src/forward/bicgstab.cpp:15-166
//   forward solve  src/forward/forward_solver_fv.cpp:84-531 (solve_pol with
//                  scattered field and post-hoc DivCorr), 546-610, 623-722
//   adjoint        :727-759, 778-888, 893-924
//   stations / Z   :1393-1485, 1490-1727
//   adjoint rhs    :1804-2013
//   sensitivity    :989-1059
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using Real = double;
using Complex = std::complex<double>;
using ComplexVec = std::vector<Complex>;
using RealVec = std::vector<Real>;
constexpr Real MU0 = 4.0e-7 * 3.14159265358979323846;
constexpr Real PI = 3.14159265358979323846;

template <typename T>
struct CSR {
    int nrows = 0, ncols = 0;
    std::vector<int> rowptr, colidx;
    std::vector<T> values;
    // common.h:52-61
    void matvec(const std::vector<T>& x, std::vector<T>& y) const {
        y.assign(nrows, T(0));
        for (int i = 0; i < nrows; ++i) {
            T sum{};
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) sum += values[k] * x[colidx[k]];
            y[i] = sum;
        }
    }
};

template <typename T>
CSR<T> from_map(int nr, int nc, const std::map<std::pair<int, int>, T>& m, bool drop_tiny) {
    CSR<T> a;
    a.nrows = nr;
    a.ncols = nc;
    a.rowptr.assign(nr + 1, 0);
    for (const auto& kv : m)
        if (!drop_tiny || std::abs(kv.second) > 1e-30) a.rowptr[kv.first.first + 1]++;
    for (int i = 0; i < nr; ++i) a.rowptr[i + 1] += a.rowptr[i];
    std::vector<int> pos(a.rowptr.begin(), a.rowptr.end() - 1);
    a.colidx.resize(a.rowptr[nr]);
    a.values.resize(a.rowptr[nr]);
    for (const auto& kv : m) {
        if (drop_tiny && !(std::abs(kv.second) > 1e-30)) continue;
        const int p = pos[kv.first.first]++;
        a.colidx[p] = kv.first.second;
        a.values[p] = kv.second;
    }
    return a;
}

struct BiCGStabResult {
    int iterations = 0;
    Real residual = 0.0;
    bool converged = false;
    int info = 0;  // 0 converged, 1 early, 2 breakdown, 3 max iterations
};

// Natural-order SSOR with omega = 1 (ssor.h:28-79).
struct SSOR {
    const CSR<Complex>* A = nullptr;
    ComplexVec inv_diag;
    mutable ComplexVec y;
    void setup(const CSR<Complex>& a) {
        A = &a;
        inv_diag.resize(a.nrows);
        y.resize(a.nrows);
        for (int i = 0; i < a.nrows; ++i) {
            Complex d(1, 0);
            for (int k = a.rowptr[i]; k < a.rowptr[i + 1]; ++k)
                if (a.colidx[k] == i) { d = a.values[k]; break; }
            inv_diag[i] = (std::abs(d) > 1e-30) ? Complex(1, 0) / d : Complex(1, 0);
        }
    }
    void apply(const ComplexVec& r, ComplexVec& out) const {
        const int n = A->nrows;
        out.resize(n);
        for (int i = 0; i < n; ++i) {
            Complex sum(0, 0);
            for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) {
                const int j = A->colidx[k];
                if (j < i) sum += A->values[k] * y[j];
            }
            y[i] = 1.0 * (r[i] - sum) * inv_diag[i];
        }
        for (int i = n - 1; i >= 0; --i) {
            Complex sum(0, 0);
            for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) {
                const int j = A->colidx[k];
                if (j > i) sum += A->values[k] * out[j];
            }
            out[i] = y[i] - 1.0 * sum * inv_diag[i];
        }
    }
};

Complex hdot(const ComplexVec& a, const ComplexVec& b) {
    Complex s(0, 0);
    for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
    return s;
}
Real norm(const ComplexVec& v) {
    Real s = 0;
    for (const auto& x : v) s += std::norm(x);
    return std::sqrt(s);
}

// Right-preconditioned BiCGStab (bicgstab.cpp:15-166).
BiCGStabResult bicgstab(const CSR<Complex>& A, const SSOR* M, const ComplexVec& b, ComplexVec& x,
                        Real tol, int maxiter) {
    BiCGStabResult res;
    const int n = static_cast<int>(b.size());
    ComplexVec r(n), r0(n), p(n), v(n), s(n), t(n), ph(n), sh(n);
    A.matvec(x, v);
    for (int i = 0; i < n; ++i) r[i] = b[i] - v[i];
    r0 = r;
    Real bnorm = norm(b);
    if (bnorm < 1e-30) bnorm = 1.0;
    Real rnorm = norm(r);
    if (rnorm / bnorm < tol) {
        res.converged = true;
        res.residual = rnorm / bnorm;
        return res;
    }
    Complex rho_old(1, 0), alpha(1, 0), omega(1, 0);
    for (int iter = 1; iter <= maxiter; ++iter) {
        const Complex rho = hdot(r0, r);
        if (std::abs(rho) < 1e-30) {
            res.iterations = iter; res.residual = norm(r) / bnorm; res.info = 2;
            return res;
        }
        if (iter == 1) {
            p = r;
        } else {
            const Complex beta = (rho / rho_old) * (alpha / omega);
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        if (M) M->apply(p, ph); else ph = p;
        A.matvec(ph, v);
        const Complex r0v = hdot(r0, v);
        if (std::abs(r0v) < 1e-30) {
            res.iterations = iter; res.residual = norm(r) / bnorm; res.info = 2;
            return res;
        }
        alpha = rho / r0v;
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        const Real snorm = norm(s);
        if (snorm / bnorm < tol) {
            for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
            res.converged = true; res.residual = snorm / bnorm; res.iterations = iter; res.info = 1;
            return res;
        }
        if (M) M->apply(s, sh); else sh = s;
        A.matvec(sh, t);
        const Complex tt = hdot(t, t);
        if (std::abs(tt) < 1e-30) {
            res.iterations = iter; res.residual = norm(s) / bnorm; res.info = 2;
            for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
            return res;
        }
        omega = hdot(t, s) / tt;
        for (int i = 0; i < n; ++i) x[i] += alpha * ph[i] + omega * sh[i];
        for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
        rnorm = norm(r);
        const Real rel = rnorm / bnorm;
        if (rel < tol) {
            res.converged = true; res.residual = rel; res.iterations = iter; res.info = 0;
            return res;
        }
        if (std::abs(omega) < 1e-30) {
            res.iterations = iter; res.residual = rel; res.info = 2;
            return res;
        }
        rho_old = rho;
    }
    res.iterations = maxiter;
    res.residual = norm(r) / bnorm;
    res.info = 3;
    return res;
}

}  // namespace

// ---------------------------------------------------------------------------
// Grid input and parameters (plain C layout for ctypes).
// ---------------------------------------------------------------------------
extern "C" {

typedef struct {
    int ne, nf, nc;
    const double *edge_x, *edge_y, *edge_z, *edge_len;
    const unsigned char* edge_dir;
    const int *edge_adj_ptr, *edge_adj;
    const double *face_x, *face_y, *face_z, *face_area;
    const unsigned char* face_normal;
    const int *face_minus, *face_plus, *face_edges_ptr, *face_edges;
    int n_hanging_faces;
    const int* hanging_faces;  // 5 per: coarse, 4 fine
    int n_hanging_edges;
    const int* hanging_edges;  // 3 per: coarse, child1, child2
    const int *cell_edges_ptr, *cell_edges;
    const double *cell_cx, *cell_cy, *cell_cz;  // mesh cell centres
    const int* cell_level;
    const unsigned char* cell_type;  // 0 earth, 1 air, 2 ocean
    double domain[6];
    double cell0_size;  // OctreeMesh::cell_size(0)
    double sigma_bg;
} orc_grid;

typedef struct {
    double bicgstab_tol;
    int bicgstab_maxiter;
    double air_z_threshold;
    int div_correction;
    int div_corr_iters;
    double div_corr_tol;
    int divcorr_outer_max;
    int divcorr_iter_per_dc;
} orc_params;

typedef struct {
    int iterations;
    double residual;
    int converged;
    int info;
    int outer;
    double true_residual;
} orc_result;

}  // extern "C"

namespace {

class Oracle {
public:
    // ---------------- operators (operators.cpp) ----------------
    void build(const orc_grid& g, const orc_params& p) {
        prm = p;
        ne = g.ne; nf = g.nf; nc = g.nc;
        dom.assign(g.domain, g.domain + 6);
        cell0_size = g.cell0_size;
        sigma_bg = g.sigma_bg > 0 ? g.sigma_bg : 0.01;
        ex.assign(g.edge_x, g.edge_x + ne); ey.assign(g.edge_y, g.edge_y + ne);
        ez.assign(g.edge_z, g.edge_z + ne); el.assign(g.edge_len, g.edge_len + ne);
        ed.assign(g.edge_dir, g.edge_dir + ne);
        adj.assign(ne, {});
        for (int e = 0; e < ne; ++e) adj[e].assign(g.edge_adj + g.edge_adj_ptr[e], g.edge_adj + g.edge_adj_ptr[e + 1]);
        fx.assign(g.face_x, g.face_x + nf); fy.assign(g.face_y, g.face_y + nf);
        fz.assign(g.face_z, g.face_z + nf); fa.assign(g.face_area, g.face_area + nf);
        fn.assign(g.face_normal, g.face_normal + nf);
        fm.assign(g.face_minus, g.face_minus + nf); fp.assign(g.face_plus, g.face_plus + nf);
        fe.assign(nf, {});
        for (int f = 0; f < nf; ++f) fe[f].assign(g.face_edges + g.face_edges_ptr[f], g.face_edges + g.face_edges_ptr[f + 1]);
        cedges.assign(nc, {});
        for (int c = 0; c < nc; ++c) cedges[c].assign(g.cell_edges + g.cell_edges_ptr[c], g.cell_edges + g.cell_edges_ptr[c + 1]);
        mcx.assign(g.cell_cx, g.cell_cx + nc); mcy.assign(g.cell_cy, g.cell_cy + nc);
        mcz.assign(g.cell_cz, g.cell_cz + nc);
        mlev.assign(g.cell_level, g.cell_level + nc);
        mtype.assign(g.cell_type, g.cell_type + nc);

        // Cell centres / half-extents from edges (:60-99).
        ccx.assign(nc, 0); ccy.assign(nc, 0); ccz.assign(nc, 0);
        hx.assign(nc, 0); hy.assign(nc, 0); hz.assign(nc, 0);
        std::vector<int> cnt(nc, 0);
        for (int e = 0; e < ne; ++e)
            for (int c : adj[e]) if (c >= 0 && c < nc) { ccx[c] += ex[e]; ccy[c] += ey[e]; ccz[c] += ez[e]; cnt[c]++; }
        for (int c = 0; c < nc; ++c)
            if (cnt[c] > 0) { ccx[c] /= cnt[c]; ccy[c] /= cnt[c]; ccz[c] /= cnt[c]; }
        for (int e = 0; e < ne; ++e) {
            const Real half = el[e] / 2.0;
            for (int c : adj[e]) {
                if (c < 0 || c >= nc) continue;
                if (ed[e] == 0 && half > hx[c]) hx[c] = half;
                if (ed[e] == 1 && half > hy[c]) hy[c] = half;
                if (ed[e] == 2 && half > hz[c]) hz[c] = half;
            }
        }
        // Face dual lengths (:103-126).
        fdl.assign(nf, 0);
        for (int f = 0; f < nf; ++f) {
            const int n = fn[f], cm = fm[f], cp = fp[f];
            auto cn = [&](int c) { return n == 0 ? ccx[c] : n == 1 ? ccy[c] : ccz[c]; };
            auto hn = [&](int c) { return n == 0 ? hx[c] : n == 1 ? hy[c] : hz[c]; };
            if (cm >= 0 && cp >= 0) fdl[f] = std::abs(cn(cp) - cn(cm));
            else {
                const int c = cm >= 0 ? cm : cp;
                fdl[f] = c >= 0 ? hn(c) : std::sqrt(fa[f]);
            }
            if (fdl[f] < 1e-20) fdl[f] = std::sqrt(fa[f]);
        }
        // Edge dual areas (:133-157).
        eda.assign(ne, 0);
        for (int e = 0; e < ne; ++e) {
            if (adj[e].empty()) { eda[e] = el[e] * el[e]; continue; }
            Real sum = 0; int count = 0;
            for (int c : adj[e]) {
                if (c < 0 || c >= nc) continue;
                Real a, b;
                if (ed[e] == 0) { a = hy[c]; b = hz[c]; }
                else if (ed[e] == 1) { a = hx[c]; b = hz[c]; }
                else { a = hx[c]; b = hy[c]; }
                sum += a * b; ++count;
            }
            eda[e] = count > 0 ? sum : el[e] * el[e];
        }
        // Curl with +-L entries (:192-244), rows in insertion order.
        std::vector<std::vector<std::pair<int, Real>>> rows(nf);
        for (int f = 0; f < nf; ++f) {
            for (int e : fe[f]) {
                if (e < 0 || e >= ne) continue;
                const int n = fn[f], d = ed[e];
                Real sign = 0;
                if (n == 0) { if (d == 1) sign = ez[e] < fz[f] ? 1 : -1; else if (d == 2) sign = ey[e] < fy[f] ? -1 : 1; }
                else if (n == 1) { if (d == 2) sign = ex[e] < fx[f] ? 1 : -1; else if (d == 0) sign = ez[e] < fz[f] ? -1 : 1; }
                else { if (d == 0) sign = ey[e] < fy[f] ? 1 : -1; else if (d == 1) sign = ex[e] < fx[f] ? -1 : 1; }
                if (std::abs(sign) > 0.5) rows[f].push_back({e, sign * el[e]});
            }
        }
        // Hanging rewrite (:344-466).
        dead_edge.assign(ne, 0);
        dead_face.assign(nf, 0);
        hang.clear();
        if (g.n_hanging_faces > 0) {
            std::map<int, std::pair<int, int>> ch;
            for (int h = 0; h < g.n_hanging_edges; ++h) {
                const int* q = g.hanging_edges + 3 * h;
                dead_edge[q[0]] = 1;
                ch[q[0]] = {q[1], q[2]};
                hang.push_back({q[0], q[1], q[2]});
            }
            for (int h = 0; h < g.n_hanging_faces; ++h) dead_face[g.hanging_faces[5 * h]] = 1;
            std::map<std::pair<int, int>, Real> merged;
            for (int f = 0; f < nf; ++f) {
                if (dead_face[f]) continue;
                for (const auto& kv : rows[f]) {
                    if (std::abs(kv.second) < 1e-30) continue;
                    if (dead_edge[kv.first]) {
                        const auto c = ch[kv.first];
                        merged[{f, c.first}] += 0.5 * kv.second;
                        merged[{f, c.second}] += 0.5 * kv.second;
                    } else {
                        merged[{f, kv.first}] += kv.second;
                    }
                }
            }
            rows.assign(nf, {});
            for (const auto& kv : merged)
                if (std::abs(kv.second) > 1e-30) rows[kv.first.first].push_back({kv.first.second, kv.second});
        }
        Ce.nrows = nf; Ce.ncols = ne;
        Ce.rowptr.assign(nf + 1, 0);
        for (int f = 0; f < nf; ++f) {
            for (const auto& kv : rows[f]) { Ce.colidx.push_back(kv.first); Ce.values.push_back(kv.second); }
            Ce.rowptr[f + 1] = static_cast<int>(Ce.colidx.size());
        }
        // Node gradient (:849-944).
        Real min_L = 1e30;
        for (int e = 0; e < ne; ++e) if (el[e] > 1e-10) min_L = std::min(min_L, el[e]);
        const Real tol = min_L * 0.01;
        std::map<std::array<long long, 3>, int> nodes;
        auto node_of = [&](Real x, Real y, Real z) {
            std::array<long long, 3> k{(long long)std::llround(x / tol), (long long)std::llround(y / tol), (long long)std::llround(z / tol)};
            auto it = nodes.find(k);
            if (it != nodes.end()) return it->second;
            const int id = static_cast<int>(nodes.size());
            nodes[k] = id;
            return id;
        };
        std::vector<int> n1(ne), n2(ne);
        for (int e = 0; e < ne; ++e) {
            Real a[3] = {ex[e], ey[e], ez[e]}, b[3] = {ex[e], ey[e], ez[e]};
            a[ed[e]] -= el[e] * 0.5;
            b[ed[e]] += el[e] * 0.5;
            n1[e] = node_of(a[0], a[1], a[2]);
            n2[e] = node_of(b[0], b[1], b[2]);
        }
        nn = static_cast<int>(nodes.size());
        G.nrows = ne; G.ncols = nn;
        G.rowptr.resize(ne + 1);
        G.colidx.resize(2 * ne);
        G.values.resize(2 * ne);
        for (int e = 0; e <= ne; ++e) G.rowptr[e] = 2 * e;
        for (int e = 0; e < ne; ++e) {
            const Real inv = 1.0 / el[e];
            if (n1[e] < n2[e]) { G.colidx[2 * e] = n1[e]; G.values[2 * e] = -inv; G.colidx[2 * e + 1] = n2[e]; G.values[2 * e + 1] = inv; }
            else { G.colidx[2 * e] = n2[e]; G.values[2 * e] = inv; G.colidx[2 * e + 1] = n1[e]; G.values[2 * e + 1] = -inv; }
        }
        // Stiffness S = Ce^T Mf^-1 Ce (slow path of :624-651 without the
        // diagonal mass), values in face order.
        RealVec mfi(nf);
        for (int f = 0; f < nf; ++f) {
            if (dead_face[f]) { mfi[f] = 0; continue; }
            const Real mf = MU0 * fa[f] / fdl[f];
            mfi[f] = mf > 1e-30 ? 1.0 / mf : 0.0;
        }
        std::map<std::pair<int, int>, Complex> smap;
        for (int f = 0; f < nf; ++f)
            for (int k1 = Ce.rowptr[f]; k1 < Ce.rowptr[f + 1]; ++k1)
                for (int k2 = Ce.rowptr[f]; k2 < Ce.rowptr[f + 1]; ++k2)
                    smap[{Ce.colidx[k2], Ce.colidx[k1]}] += Complex(Ce.values[k2] * mfi[f] * Ce.values[k1], 0.0);
        for (int e = 0; e < ne; ++e) smap[{e, e}] += Complex(0, 0);
        S = from_map<Complex>(ne, ne, smap, false);

        // Solve Dirichlet set (forward_solver_fv.cpp:117, 166-194).
        const Real bt = cell0_size * 0.1;
        is_bc.assign(ne, 0);
        for (int e = 0; e < ne; ++e) {
            const bool on = ex[e] <= dom[0] + bt || ex[e] >= dom[1] - bt || ey[e] <= dom[2] + bt ||
                            ey[e] >= dom[3] - bt || ez[e] <= dom[4] + bt || ez[e] >= dom[5] - bt;
            if (on || ez[e] > prm.air_z_threshold) is_bc[e] = 1;
        }
        is_dead.assign(ne, 0);
        for (const auto& h : hang) is_dead[h[0]] = 1;
        // DivCorr Dirichlet set (:556-579).
        dc_bc.assign(ne, 0);
        for (int e = 0; e < ne; ++e) {
            if (dead_edge[e]) continue;
            const Real t = 1.0;
            if (ex[e] <= dom[0] + t || ex[e] >= dom[1] - t || ey[e] <= dom[2] + t || ey[e] >= dom[3] - t ||
                ez[e] <= dom[4] + t || ez[e] >= dom[5] - t || ez[e] > prm.air_z_threshold)
                dc_bc[e] = 1;
        }
        dc_bc_node.assign(nn, 0);
        for (int e = 0; e < ne; ++e)
            if (dc_bc[e]) for (int k = G.rowptr[e]; k < G.rowptr[e + 1]; ++k) dc_bc_node[G.colidx[k]] = 1;
    }

    void set_sigma(const RealVec& s) {
        sigma = s;
        // Me (:481-518) and Me_sigma (:956-973).
        me.assign(ne, 0);
        mes.assign(ne, 0);
        for (int e = 0; e < ne; ++e) {
            if (adj[e].empty()) continue;
            Real sum = 0; int count = 0;
            for (int c : adj[e]) if (c >= 0 && c < (int)sigma.size()) { sum += sigma[c]; ++count; }
            const Real avg = count > 0 ? sum / count : 1e-8;
            me[e] = avg * (eda[e] * el[e]);
            if (!dead_edge[e]) mes[e] = avg * eda[e] * el[e];
        }
        for (const auto& h : hang) me[h[0]] = 0;
        // L = G^T Me_sigma G (:976-1015) with the Dirichlet rewrite (:584-606).
        std::map<std::pair<int, int>, Real> lm;
        for (int e = 0; e < ne; ++e) {
            const Real w = mes[e];
            if (w < 1e-30) continue;
            for (int k1 = G.rowptr[e]; k1 < G.rowptr[e + 1]; ++k1)
                for (int k2 = G.rowptr[e]; k2 < G.rowptr[e + 1]; ++k2)
                    lm[{G.colidx[k1], G.colidx[k2]}] += G.values[k1] * w * G.values[k2];
        }
        L = from_map<Real>(nn, nn, lm, false);
        Ldiag.assign(nn, 1.0);
        for (int i = 0; i < nn; ++i) {
            if (dc_bc_node[i]) continue;
            for (int k = L.rowptr[i]; k < L.rowptr[i + 1]; ++k) if (L.colidx[k] == i) { Ldiag[i] = L.values[k]; break; }
        }
        for (int i = 0; i < nn; ++i) {
            if (!dc_bc_node[i]) continue;
            for (int k = L.rowptr[i]; k < L.rowptr[i + 1]; ++k) L.values[k] = L.colidx[k] == i ? 1.0 : 0.0;
        }
        for (int i = 0; i < nn; ++i) {
            if (dc_bc_node[i]) continue;
            for (int k = L.rowptr[i]; k < L.rowptr[i + 1]; ++k) if (dc_bc_node[L.colidx[k]]) L.values[k] = 0.0;
        }
    }

    // A = S + i omega Me, then identity elimination of BC + dead rows / cols
    // (:224-235); the saved original entries feed the right-hand side.
    void assemble(Real omega) {
        w = omega;
        A = S;
        const Complex iw(0.0, omega);
        for (int e = 0; e < ne; ++e)
            for (int k = A.rowptr[e]; k < A.rowptr[e + 1]; ++k)
                if (A.colidx[k] == e) A.values[k] += iw * Complex(me[e], 0.0);
        Aorig = A;
        std::vector<char> cons(ne);
        for (int i = 0; i < ne; ++i) cons[i] = is_bc[i] || is_dead[i];
        for (int i = 0; i < ne; ++i)
            for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) {
                const int j = A.colidx[k];
                if (cons[i] || cons[j]) A.values[k] = cons[i] ? (j == i ? Complex(1, 0) : Complex(0, 0)) : Complex(0, 0);
            }
        constrained = cons;
        ssor.setup(A);
    }

    Complex E_primary(Real z) const {
        const Complex ik = std::sqrt(Complex(0, w * MU0 * sigma_bg));
        if (z >= 0.0) return Complex(1.0, 0.0) + ik * z;
        return std::exp(ik * z);
    }

    ComplexVec build_rhs(const ComplexVec& ebc) const {
        ComplexVec rhs(ne, Complex(0, 0));
        for (int i = 0; i < ne; ++i) {
            if (constrained[i]) continue;
            for (int k = Aorig.rowptr[i]; k < Aorig.rowptr[i + 1]; ++k) {
                const int j = Aorig.colidx[k];
                if (is_bc[j]) rhs[i] -= Aorig.values[k] * ebc[j];
            }
        }
        for (int i = 0; i < ne; ++i) if (is_bc[i]) rhs[i] = ebc[i];
        return rhs;
    }

    ComplexVec bc_values(int dir) const {
        ComplexVec v(ne, Complex(0, 0));
        for (int e = 0; e < ne; ++e)
            if (is_bc[e] && ed[e] == dir) v[e] = E_primary(ez[e]);
        return v;
    }

    void interpolate_hanging(ComplexVec& E) const {
        for (const auto& h : hang) E[h[0]] = Complex(0.5) * (E[h[1]] + E[h[2]]);
    }

    // divcorr_poisson_ (:623-722).
    void divcorr(ComplexVec& E, const ComplexVec* phi0) const {
        Real dn0 = -1;
        for (int it = 0; it < prm.div_corr_iters; ++it) {
            ComplexVec divJ(nn, Complex(0, 0));
            for (int e = 0; e < ne; ++e) {
                if (dead_edge[e]) continue;
                const Complex we = mes[e] * E[e];
                for (int k = G.rowptr[e]; k < G.rowptr[e + 1]; ++k) divJ[G.colidx[k]] += G.values[k] * we;
            }
            if (phi0) for (int i = 0; i < nn; ++i) divJ[i] -= (*phi0)[i];
            for (int i = 0; i < nn; ++i) if (dc_bc_node[i]) divJ[i] = Complex(0, 0);
            Real dn = 0;
            for (int i = 0; i < nn; ++i) dn += std::norm(divJ[i]);
            dn = std::sqrt(dn);
            if (it == 0) dn0 = dn;
            const Real rel = dn0 > 1e-30 ? dn / dn0 : 0;
            if ((rel < prm.div_corr_tol && it > 0) || dn < 1e-12) break;
            ComplexVec phi(nn, Complex(0, 0)), r = divJ, z(nn), p(nn), Ap(nn);
            for (int i = 0; i < nn; ++i) z[i] = std::abs(Ldiag[i]) > 1e-30 ? r[i] / Ldiag[i] : r[i];
            p = z;
            Complex rz(0, 0);
            for (int i = 0; i < nn; ++i) rz += std::conj(r[i]) * z[i];
            const int cg_max = std::max(500, nn);
            for (int cg = 0; cg < cg_max; ++cg) {
                for (int i = 0; i < nn; ++i) {
                    Ap[i] = Complex(0, 0);
                    for (int k = L.rowptr[i]; k < L.rowptr[i + 1]; ++k) Ap[i] += L.values[k] * p[L.colidx[k]];
                }
                Complex pAp(0, 0);
                for (int i = 0; i < nn; ++i) pAp += std::conj(p[i]) * Ap[i];
                if (std::abs(pAp) < 1e-30) break;
                const Complex alpha = rz / pAp;
                for (int i = 0; i < nn; ++i) { phi[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
                Real rn = 0;
                for (int i = 0; i < nn; ++i) rn += std::norm(r[i]);
                if (std::sqrt(rn) < 1e-10 * dn) break;
                for (int i = 0; i < nn; ++i) z[i] = std::abs(Ldiag[i]) > 1e-30 ? r[i] / Ldiag[i] : r[i];
                Complex rzn(0, 0);
                for (int i = 0; i < nn; ++i) rzn += std::conj(r[i]) * z[i];
                const Complex beta = rzn / (rz + Complex(1e-30, 0));
                for (int i = 0; i < nn; ++i) p[i] = z[i] + beta * p[i];
                rz = rzn;
            }
            for (int e = 0; e < ne; ++e) {
                Complex gp(0, 0);
                for (int k = G.rowptr[e]; k < G.rowptr[e + 1]; ++k) gp += G.values[k] * phi[G.colidx[k]];
                if (!dc_bc[e]) E[e] -= gp;
            }
        }
    }

    // solve_pol with the scattered-field formulation (:301-367).
    BiCGStabResult solve_pol(int dir, ComplexVec& E) const {
        const ComplexVec ebc = bc_values(dir);
        const ComplexVec rhs = build_rhs(ebc);
        ComplexVec epri(ne, Complex(0, 0));
        for (int e = 0; e < ne; ++e) if (ed[e] == dir) epri[e] = E_primary(ez[e]);
        ComplexVec aep;
        A.matvec(epri, aep);
        ComplexVec rs(ne);
        for (int i = 0; i < ne; ++i) rs[i] = rhs[i] - aep[i];
        ComplexVec es(ne, Complex(0, 0));
        const BiCGStabResult r = bicgstab(A, &ssor, rs, es, prm.bicgstab_tol, prm.bicgstab_maxiter);
        if (!hang.empty()) interpolate_hanging(es);
        E.assign(ne, Complex(0, 0));
        for (int i = 0; i < ne; ++i) E[i] = epri[i] + es[i];
        for (int i = 0; i < ne; ++i) if (is_bc[i]) E[i] = ebc[i];
        if (prm.div_correction) divcorr(E, nullptr);
        return r;
    }

    Real true_residual(const ComplexVec& x, const ComplexVec& b) const {
        ComplexVec ax;
        A.matvec(x, ax);
        Real rn = 0, bn = 0;
        for (int i = 0; i < ne; ++i) { rn += std::norm(ax[i] - b[i]); bn += std::norm(b[i]); }
        bn = std::sqrt(bn);
        if (bn < 1e-30) bn = 1.0;
        return std::sqrt(rn) / bn;
    }

    // adjoint_solve / solve_rhs (:914-940) via solve_with_divcorr_ (:778-888)
    // or solve_system_ (:727-759).
    orc_result adjoint(const ComplexVec& rhs, ComplexVec& x) const {
        orc_result out{};
        x.assign(ne, Complex(0, 0));
        if (prm.div_correction && w > 0) {
            ComplexVec phi0(nn, Complex(0, 0));
            for (int e = 0; e < ne; ++e)
                if (!dead_edge[e]) for (int k = G.rowptr[e]; k < G.rowptr[e + 1]; ++k) phi0[G.colidx[k]] += G.values[k] * rhs[e];
            const Complex iw(0, w);
            for (auto& v : phi0) v /= iw;
            Real rn = norm(rhs);
            if (rn < 1e-30) rn = 1.0;
            divcorr(x, &phi0);
            Real prev = 1e30;
            for (int outer = 0; outer < prm.divcorr_outer_max; ++outer) {
                const BiCGStabResult r = bicgstab(A, &ssor, rhs, x, prm.bicgstab_tol, prm.divcorr_iter_per_dc);
                out.iterations += r.iterations;
                out.residual = r.residual;
                out.info = r.info;
                out.outer = outer + 1;
                if (r.residual > 1.0) {
                    std::fill(x.begin(), x.end(), Complex(0, 0));
                    break;
                }
                if (r.converged) { out.converged = 1; break; }
                divcorr(x, &phi0);
                ComplexVec ax;
                A.matvec(x, ax);
                Real res = 0;
                for (int i = 0; i < ne; ++i) res += std::norm(ax[i] - rhs[i]);
                const Real rel = std::sqrt(res) / rn;
                if (rel < prm.bicgstab_tol) { out.converged = 1; break; }
                if (outer > 0 && rel > 0.95 * prev) break;
                prev = rel;
            }
        } else {
            const BiCGStabResult r = bicgstab(A, &ssor, rhs, x, prm.bicgstab_tol, prm.bicgstab_maxiter);
            out.iterations = r.iterations; out.residual = r.residual; out.info = r.info;
            out.converged = r.converged; out.outer = 1;
            if (!r.converged && r.residual > 1.0) std::fill(x.begin(), x.end(), Complex(0, 0));
        }
        out.true_residual = true_residual(x, rhs);
        return out;
    }

    // find_stations_ (:1393-1485) from the mesh cell data.
    void find_stations(const double* xy, int n) {
        station_cell.clear();
        for (int s = 0; s < n; ++s) {
            const Real sx = xy[2 * s], sy = xy[2 * s + 1], sz = 0.0;
            int best = -1, fb = -1;
            Real bcz = 1e30, fd = 1e30;
            for (int c = 0; c < nc; ++c) {
                if (mtype[c] == 1) continue;
                const Real hu = 1.0 / (1 << mlev[c]);
                const Real hdx = hu * (dom[1] - dom[0]), hdy = hu * (dom[3] - dom[2]);
                const bool inx = sx >= mcx[c] - hdx / 2 - 1.0 && sx <= mcx[c] + hdx / 2 + 1.0;
                const bool iny = sy >= mcy[c] - hdy / 2 - 1.0 && sy <= mcy[c] + hdy / 2 + 1.0;
                if (inx && iny && mcz[c] < 0 && std::abs(mcz[c]) < bcz) { bcz = std::abs(mcz[c]); best = c; }
                const Real d = std::sqrt((mcx[c] - sx) * (mcx[c] - sx) + (mcy[c] - sy) * (mcy[c] - sy) + (mcz[c] - sz) * (mcz[c] - sz));
                if (d < fd) { fd = d; fb = c; }
            }
            station_cell.push_back(best >= 0 ? best : fb);
        }
    }

    struct HC { Complex Hx1, Hy1, Hx2, Hy2, det, Zxx, Zxy, Zyx, Zyy; Real cdx, cdy, cdz; };

    // extract_impedance_ (:1490-1727).
    void impedance(const ComplexVec& E1, const ComplexVec& E2, std::vector<std::array<Complex, 4>>& Z) {
        const Complex iwmu(0, w * MU0);
        hcache.assign(station_cell.size(), HC{});
        Z.assign(station_cell.size(), {Complex(0, 0), Complex(0, 0), Complex(0, 0), Complex(0, 0)});
        for (size_t s = 0; s < station_cell.size(); ++s) {
            const int c = station_cell[s];
            if (c < 0) continue;
            const Real cx = mcx[c], cy = mcy[c], cz = mcz[c];
            const Real hu = 1.0 / (1 << mlev[c]);
            const Real cdx = hu * (dom[1] - dom[0]), cdy = hu * (dom[3] - dom[2]), cdz = hu * (dom[5] - dom[4]);
            Complex x1t, x1b, x2t, x2b, y1t, y1b, y2t, y2b, z1px, z1mx, z2px, z2mx, z1py, z1my, z2py, z2my;
            int nxt = 0, nxb = 0, nyt = 0, nyb = 0, npx = 0, nmx = 0, npy = 0, nmy = 0;
            for (int e : cedges[c]) {
                if (ed[e] == 0) { if (ez[e] > cz) { x1t += E1[e]; x2t += E2[e]; ++nxt; } else { x1b += E1[e]; x2b += E2[e]; ++nxb; } }
                else if (ed[e] == 1) { if (ez[e] > cz) { y1t += E1[e]; y2t += E2[e]; ++nyt; } else { y1b += E1[e]; y2b += E2[e]; ++nyb; } }
                else {
                    if (ex[e] > cx) { z1px += E1[e]; z2px += E2[e]; ++npx; } else { z1mx += E1[e]; z2mx += E2[e]; ++nmx; }
                    if (ey[e] > cy) { z1py += E1[e]; z2py += E2[e]; ++npy; } else { z1my += E1[e]; z2my += E2[e]; ++nmy; }
                }
            }
            if (nxt > 0) { x1t /= Complex(nxt, 0); x2t /= Complex(nxt, 0); }
            if (nxb > 0) { x1b /= Complex(nxb, 0); x2b /= Complex(nxb, 0); }
            if (nyt > 0) { y1t /= Complex(nyt, 0); y2t /= Complex(nyt, 0); }
            if (nyb > 0) { y1b /= Complex(nyb, 0); y2b /= Complex(nyb, 0); }
            if (npx > 0) { z1px /= Complex(npx, 0); z2px /= Complex(npx, 0); }
            if (nmx > 0) { z1mx /= Complex(nmx, 0); z2mx /= Complex(nmx, 0); }
            if (npy > 0) { z1py /= Complex(npy, 0); z2py /= Complex(npy, 0); }
            if (nmy > 0) { z1my /= Complex(nmy, 0); z2my /= Complex(nmy, 0); }
            const Complex Ex1 = (nxt > 0 && nxb > 0) ? (x1t + x1b) * Complex(0.5, 0) : (nxt > 0 ? x1t : x1b);
            const Complex Ey1 = (nyt > 0 && nyb > 0) ? (y1t + y1b) * Complex(0.5, 0) : (nyt > 0 ? y1t : y1b);
            const Complex Ex2 = (nxt > 0 && nxb > 0) ? (x2t + x2b) * Complex(0.5, 0) : (nxt > 0 ? x2t : x2b);
            const Complex Ey2 = (nyt > 0 && nyb > 0) ? (y2t + y2b) * Complex(0.5, 0) : (nyt > 0 ? y2t : y2b);
            Complex a1, a2, b1, b2, c1, c2, d1, d2;  // dEx/dz, dEy/dz, dEz/dx, dEz/dy per pol
            if (nxt > 0 && nxb > 0) { a1 = (x1t - x1b) / cdz; a2 = (x2t - x2b) / cdz; }
            if (nyt > 0 && nyb > 0) { b1 = (y1t - y1b) / cdz; b2 = (y2t - y2b) / cdz; }
            if (npx > 0 && nmx > 0) { c1 = (z1px - z1mx) / cdx; c2 = (z2px - z2mx) / cdx; }
            if (npy > 0 && nmy > 0) { d1 = (z1py - z1my) / cdy; d2 = (z2py - z2my) / cdy; }
            HC& h = hcache[s];
            h.Hx1 = (d1 - b1) / iwmu; h.Hy1 = (a1 - c1) / iwmu;
            h.Hx2 = (d2 - b2) / iwmu; h.Hy2 = (a2 - c2) / iwmu;
            h.det = h.Hx1 * h.Hy2 - h.Hx2 * h.Hy1;
            h.cdx = cdx; h.cdy = cdy; h.cdz = cdz;
            if (std::abs(h.det) > 1e-30) {
                h.Zxx = (Ex1 * h.Hy2 - Ex2 * h.Hy1) / h.det;
                h.Zxy = (Ex2 * h.Hx1 - Ex1 * h.Hx2) / h.det;
                h.Zyx = (Ey1 * h.Hy2 - Ey2 * h.Hy1) / h.det;
                h.Zyy = (Ey2 * h.Hx1 - Ey1 * h.Hx2) / h.det;
            } else {
                const Complex z1d = std::sqrt(iwmu / Complex(sigma[c], 0));
                h.Zxx = 0; h.Zxy = z1d; h.Zyx = -z1d; h.Zyy = 0;
            }
            Z[s] = {h.Zxx, h.Zxy, h.Zyx, h.Zyy};
        }
    }

    // build_adjoint_rhs_from_residual (:1804-2013).
    void adjoint_rhs(const std::vector<std::array<Complex, 4>>& wr, ComplexVec& r1, ComplexVec& r2) const {
        r1.assign(ne, Complex(0, 0));
        r2.assign(ne, Complex(0, 0));
        const Complex iwmu(0, w * MU0);
        for (size_t s = 0; s < station_cell.size(); ++s) {
            const int c = station_cell[s];
            if (c < 0) continue;
            const HC& h = hcache[s];
            if (std::abs(h.det) < 1e-30) continue;
            const auto& q = wr[s];
            const Complex sx1 = (q[0] * h.Hy2 - q[1] * h.Hx2) / h.det;
            const Complex sy1 = (q[2] * h.Hy2 - q[3] * h.Hx2) / h.det;
            const Complex sx2 = (-q[0] * h.Hy1 + q[1] * h.Hx1) / h.det;
            const Complex sy2 = (-q[2] * h.Hy1 + q[3] * h.Hx1) / h.det;
            const Complex hx1 = -(h.Zxx * sx1 + h.Zyx * sy1), hy1 = -(h.Zxy * sx1 + h.Zyy * sy1);
            const Complex hx2 = -(h.Zxx * sx2 + h.Zyx * sy2), hy2 = -(h.Zxy * sx2 + h.Zyy * sy2);
            const Real cx = mcx[c], cy = mcy[c], cz = mcz[c];
            int nxt = 0, nxb = 0, nyt = 0, nyb = 0, npx = 0, nmx = 0, npy = 0, nmy = 0;
            for (int e : cedges[c]) {
                if (ed[e] == 0) { if (ez[e] > cz) ++nxt; else ++nxb; }
                else if (ed[e] == 1) { if (ez[e] > cz) ++nyt; else ++nyb; }
                else { if (ex[e] > cx) ++npx; else ++nmx; if (ey[e] > cy) ++npy; else ++nmy; }
            }
            Real wxt = nxt > 0 ? 0.5 / nxt : 0, wxb = nxb > 0 ? 0.5 / nxb : 0;
            Real wyt = nyt > 0 ? 0.5 / nyt : 0, wyb = nyb > 0 ? 0.5 / nyb : 0;
            if (nxt == 0) wxb = nxb > 0 ? 1.0 / nxb : 0;
            if (nxb == 0) wxt = nxt > 0 ? 1.0 / nxt : 0;
            if (nyt == 0) wyb = nyb > 0 ? 1.0 / nyb : 0;
            if (nyb == 0) wyt = nyt > 0 ? 1.0 / nyt : 0;
            const Complex iz = Complex(1, 0) / (iwmu * h.cdz);
            const Complex ix = Complex(1, 0) / (iwmu * h.cdx);
            const Complex iy = Complex(1, 0) / (iwmu * h.cdy);
            for (int e : cedges[c]) {
                if (ed[e] == 0) {
                    const bool top = ez[e] > cz;
                    const Real we = top ? wxt : wxb;
                    r1[e] += sx1 * we; r2[e] += sx2 * we;
                    if (top && nxt > 0) { const Complex wh = iz / Real(nxt); r1[e] += hy1 * wh; r2[e] += hy2 * wh; }
                    else if (!top && nxb > 0) { const Complex wh = -iz / Real(nxb); r1[e] += hy1 * wh; r2[e] += hy2 * wh; }
                } else if (ed[e] == 1) {
                    const bool top = ez[e] > cz;
                    const Real we = top ? wyt : wyb;
                    r1[e] += sy1 * we; r2[e] += sy2 * we;
                    if (top && nyt > 0) { const Complex wh = -iz / Real(nyt); r1[e] += hx1 * wh; r2[e] += hx2 * wh; }
                    else if (!top && nyb > 0) { const Complex wh = iz / Real(nyb); r1[e] += hx1 * wh; r2[e] += hx2 * wh; }
                } else {
                    const bool py = ey[e] > cy;
                    if (py && npy > 0) { const Complex wh = iy / Real(npy); r1[e] += hx1 * wh; r2[e] += hx2 * wh; }
                    else if (!py && nmy > 0) { const Complex wh = -iy / Real(nmy); r1[e] += hx1 * wh; r2[e] += hx2 * wh; }
                    const bool px = ex[e] > cx;
                    if (px && npx > 0) { const Complex wh = -ix / Real(npx); r1[e] += hy1 * wh; r2[e] += hy2 * wh; }
                    else if (!px && nmx > 0) { const Complex wh = ix / Real(nmx); r1[e] += hy1 * wh; r2[e] += hy2 * wh; }
                }
            }
        }
    }

    // compute_sensitivity (:989-1059).
    void sensitivity(const ComplexVec& Eb, const ComplexVec& lam, RealVec& g) const {
        g.assign(nc, 0.0);
        std::vector<char> cons(ne, 0);
        for (int e = 0; e < ne; ++e) {
            if (dead_edge[e]) { cons[e] = 1; continue; }
            const Real t = 1.0;
            if (ex[e] <= dom[0] + t || ex[e] >= dom[1] - t || ey[e] <= dom[2] + t || ey[e] >= dom[3] - t ||
                ez[e] <= dom[4] + t || ez[e] >= dom[5] - t || ez[e] > prm.air_z_threshold)
                cons[e] = 1;
        }
        const Complex iw(0, w);
        for (int c = 0; c < nc; ++c) {
            Complex s(0, 0);
            for (int e : cedges[c]) {
                if (cons[e]) continue;
                int na = 0;
                for (int a : adj[e]) if (a >= 0 && a < nc) ++na;
                if (na == 0) continue;
                s += lam[e] * Eb[e] * (sigma[c] * eda[e] * el[e] / na);
            }
            g[c] = (iw * s).real();
        }
    }

    // build_perturbation_rhs (:945-984): delta b = -i w sum_c dMe/dsigma_c
    // delta_sigma_c E_bg with the reference's harmonic-mean, L^3 mass
    // derivative, cells in order, each cell's edges in cell-edge order.
    void perturbation_rhs(const ComplexVec& Eb, const RealVec& dsig, ComplexVec& rhs) const {
        rhs.assign(ne, Complex(0, 0));
        const Complex iw(0, w);
        for (int c = 0; c < nc; ++c) {
            if (std::abs(dsig[c]) < 1e-30) continue;
            const Real sc = sigma[c];
            for (int e : cedges[c]) {
                const Real L = el[e], vol = L * L * L;
                Real sum_inv = 0.0;
                int na = 0;
                for (int a : adj[e])
                    if (a >= 0 && a < nc && sigma[a] > 1e-30) { sum_inv += 1.0 / sigma[a]; ++na; }
                const Real savg = (na > 0 && sum_inv > 1e-30) ? na / sum_inv : 1e-8;
                const Real dme = vol * savg * savg / (na * sc * sc);
                rhs[e] -= iw * Complex(dsig[c] * dme, 0) * Eb[e];
            }
        }
    }

    // extract_delta_impedance_complex (:2018-2065): station averages of the
    // x and y components of dE over all the cell's x / y edges, times the
    // column of H^-1 of the polarisation (1 or 2), H from the cache of the
    // last impedance call; delta H is not formed.
    void delta_impedance(const ComplexVec& dE, int pol, std::vector<std::array<Complex, 4>>& dz) const {
        dz.assign(station_cell.size(), {Complex(0, 0), Complex(0, 0), Complex(0, 0), Complex(0, 0)});
        for (size_t s = 0; s < station_cell.size(); ++s) {
            const int c = station_cell[s];
            if (c < 0) continue;
            Complex dx(0, 0), dy(0, 0);
            int nx = 0, ny = 0;
            for (int e : cedges[c]) {
                if (ed[e] == 0) { dx += dE[e]; ++nx; }
                else if (ed[e] == 1) { dy += dE[e]; ++ny; }
            }
            if (nx > 0) dx /= Complex(nx, 0);
            if (ny > 0) dy /= Complex(ny, 0);
            const HC& h = hcache[s];
            if (std::abs(h.det) < 1e-30) continue;
            const Complex inv = Complex(1, 0) / h.det;
            if (pol == 1) dz[s] = {dx * h.Hy2 * inv, dx * (-h.Hx2) * inv, dy * h.Hy2 * inv, dy * (-h.Hx2) * inv};
            else dz[s] = {dx * (-h.Hy1) * inv, dx * h.Hx1 * inv, dy * (-h.Hy1) * inv, dy * h.Hx1 * inv};
        }
    }

    orc_params prm{};
    int ne = 0, nf = 0, nc = 0, nn = 0;
    std::vector<double> dom;
    Real cell0_size = 0, sigma_bg = 0.01, w = 0;
    RealVec ex, ey, ez, el, fx, fy, fz, fa, mcx, mcy, mcz;
    std::vector<int> ed, fn, fm, fp, mlev, station_cell;
    std::vector<unsigned char> mtype;
    std::vector<std::vector<int>> adj, fe, cedges;
    RealVec ccx, ccy, ccz, hx, hy, hz, fdl, eda;
    std::vector<char> dead_edge, dead_face, is_bc, is_dead, dc_bc, dc_bc_node, constrained;
    std::vector<std::array<int, 3>> hang;
    CSR<Real> Ce, G, L;
    CSR<Complex> S, A, Aorig;
    RealVec sigma, me, mes, Ldiag;
    SSOR ssor;
    std::vector<HC> hcache;
};

thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
    try { f(); return 0; } catch (const std::exception& e) { g_err = e.what(); return 1; }
}

const Complex* cptr(const double* p) { return reinterpret_cast<const Complex*>(p); }

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_create(const orc_grid* g, const orc_params* p) {
    Oracle* o = new Oracle;
    if (guard([&] { o->build(*g, *p); }) != 0) { delete o; return nullptr; }
    return o;
}
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }
int orc_num_nodes(void* h) { return static_cast<Oracle*>(h)->nn; }

int orc_set_sigma(void* h, const double* sigma) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] { o->set_sigma(RealVec(sigma, sigma + o->nc)); });
}

int orc_assemble(void* h, double omega) {
    return guard([&] { static_cast<Oracle*>(h)->assemble(omega); });
}

// Named arrays of the oracle: "S" (nnz complex + pattern), "A", "G", "L", etc.
int orc_get_csr(void* h, const char* name, int* nrows, int* nnz, const int** ptr, const int** col,
                const double** val, int* is_complex) {
    Oracle* o = static_cast<Oracle*>(h);
    const std::string n(name);
    return guard([&] {
        if (n == "A" || n == "S" || n == "Aorig") {
            const CSR<Complex>& m = n == "A" ? o->A : n == "S" ? o->S : o->Aorig;
            *nrows = m.nrows; *nnz = (int)m.colidx.size(); *ptr = m.rowptr.data(); *col = m.colidx.data();
            *val = reinterpret_cast<const double*>(m.values.data()); *is_complex = 1;
        } else {
            const CSR<Real>& m = n == "Ce" ? o->Ce : n == "G" ? o->G : o->L;
            *nrows = m.nrows; *nnz = (int)m.colidx.size(); *ptr = m.rowptr.data(); *col = m.colidx.data();
            *val = m.values.data(); *is_complex = 0;
        }
    });
}

int orc_get_vec(void* h, const char* name, const double** data, int* count) {
    Oracle* o = static_cast<Oracle*>(h);
    const std::string n(name);
    return guard([&] {
        const RealVec* v = n == "me" ? &o->me : n == "mes" ? &o->mes : n == "Ldiag" ? &o->Ldiag :
                           n == "eda" ? &o->eda : n == "fdl" ? &o->fdl : n == "ccx" ? &o->ccx :
                           n == "ccy" ? &o->ccy : n == "ccz" ? &o->ccz : nullptr;
        if (!v) throw std::invalid_argument("unknown vector " + n);
        *data = v->data();
        *count = (int)v->size();
    });
}

int orc_spmv(void* h, const double* x, double* y) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec xv(cptr(x), cptr(x) + o->ne), yv;
        o->A.matvec(xv, yv);
        std::memcpy(y, yv.data(), yv.size() * sizeof(Complex));
    });
}

int orc_ssor(void* h, const double* r, double* z) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec rv(cptr(r), cptr(r) + o->ne), zv;
        o->ssor.apply(rv, zv);
        std::memcpy(z, zv.data(), zv.size() * sizeof(Complex));
    });
}

// Generic SSOR-BiCGStab on a caller CSR (for permuted-order equivalence tests).
int orc_bicgstab_csr(int n, const int* ptr, const int* col, const double* val, const double* b,
                     double* x, double tol, int maxiter, int use_ssor, orc_result* res) {
    return guard([&] {
        CSR<Complex> A;
        A.nrows = A.ncols = n;
        A.rowptr.assign(ptr, ptr + n + 1);
        A.colidx.assign(col, col + ptr[n]);
        A.values.assign(cptr(val), cptr(val) + ptr[n]);
        SSOR M;
        M.setup(A);
        ComplexVec bv(cptr(b), cptr(b) + n), xv(cptr(x), cptr(x) + n);
        const BiCGStabResult r = bicgstab(A, use_ssor ? &M : nullptr, bv, xv, tol, maxiter);
        std::memcpy(x, xv.data(), n * sizeof(Complex));
        res->iterations = r.iterations; res->residual = r.residual; res->converged = r.converged;
        res->info = r.info; res->outer = 1;
    });
}

int orc_ssor_csr(int n, const int* ptr, const int* col, const double* val, const double* r, double* z) {
    return guard([&] {
        CSR<Complex> A;
        A.nrows = A.ncols = n;
        A.rowptr.assign(ptr, ptr + n + 1);
        A.colidx.assign(col, col + ptr[n]);
        A.values.assign(cptr(val), cptr(val) + ptr[n]);
        SSOR M;
        M.setup(A);
        ComplexVec rv(cptr(r), cptr(r) + n), zv;
        M.apply(rv, zv);
        std::memcpy(z, zv.data(), n * sizeof(Complex));
    });
}

int orc_divcorr(void* h, double* E, const double* phi0) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec ev(cptr(E), cptr(E) + o->ne);
        ComplexVec p;
        if (phi0) p.assign(cptr(phi0), cptr(phi0) + o->nn);
        o->divcorr(ev, phi0 ? &p : nullptr);
        std::memcpy(E, ev.data(), ev.size() * sizeof(Complex));
    });
}

// One polarisation of solve_pol (dir 0: Ex, 1: Ey); call orc_assemble first.
int orc_solve_pol(void* h, int dir, double* E, orc_result* res) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec ev;
        const BiCGStabResult r = o->solve_pol(dir, ev);
        std::memcpy(E, ev.data(), ev.size() * sizeof(Complex));
        res->iterations = r.iterations; res->residual = r.residual; res->converged = r.converged;
        res->info = r.info; res->outer = 1; res->true_residual = 0;
    });
}

int orc_rhs(void* h, int dir, double* rhs_total, double* epri) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        const ComplexVec r = o->build_rhs(o->bc_values(dir));
        std::memcpy(rhs_total, r.data(), r.size() * sizeof(Complex));
        for (int e = 0; e < o->ne; ++e) {
            const Complex v = o->ed[e] == dir ? o->E_primary(o->ez[e]) : Complex(0, 0);
            std::memcpy(epri + 2 * e, &v, sizeof(Complex));
        }
    });
}

int orc_adjoint(void* h, const double* rhs, double* x, orc_result* res) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec rv(cptr(rhs), cptr(rhs) + o->ne), xv;
        *res = o->adjoint(rv, xv);
        std::memcpy(x, xv.data(), xv.size() * sizeof(Complex));
    });
}

int orc_find_stations(void* h, const double* xy, int n, int* cells) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        o->find_stations(xy, n);
        std::copy(o->station_cell.begin(), o->station_cell.end(), cells);
    });
}

int orc_impedance(void* h, const double* e1, const double* e2, double* z) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec a(cptr(e1), cptr(e1) + o->ne), b(cptr(e2), cptr(e2) + o->ne);
        std::vector<std::array<Complex, 4>> Z;
        o->impedance(a, b, Z);
        for (size_t s = 0; s < Z.size(); ++s)
            for (int k = 0; k < 4; ++k) { z[8 * s + 2 * k] = Z[s][k].real(); z[8 * s + 2 * k + 1] = Z[s][k].imag(); }
    });
}

int orc_adjoint_rhs(void* h, const double* wr, double* r1, double* r2) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        std::vector<std::array<Complex, 4>> w(o->station_cell.size());
        for (size_t s = 0; s < w.size(); ++s)
            for (int k = 0; k < 4; ++k) w[s][k] = Complex(wr[8 * s + 2 * k], wr[8 * s + 2 * k + 1]);
        ComplexVec a, b;
        o->adjoint_rhs(w, a, b);
        std::memcpy(r1, a.data(), a.size() * sizeof(Complex));
        std::memcpy(r2, b.data(), b.size() * sizeof(Complex));
    });
}

int orc_perturbation_rhs(void* h, const double* ebg, const double* dsigma, double* rhs) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec a(cptr(ebg), cptr(ebg) + o->ne), r;
        RealVec d(dsigma, dsigma + o->nc);
        o->perturbation_rhs(a, d, r);
        std::copy(r.begin(), r.end(), reinterpret_cast<Complex*>(rhs));
    });
}

int orc_delta_impedance(void* h, const double* de, int pol, double* dz) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        if (pol != 1 && pol != 2) throw std::invalid_argument("polarisation is 1 or 2");
        ComplexVec a(cptr(de), cptr(de) + o->ne);
        std::vector<std::array<Complex, 4>> z;
        o->delta_impedance(a, pol, z);
        Complex* out = reinterpret_cast<Complex*>(dz);
        for (size_t s = 0; s < z.size(); ++s)
            for (int k = 0; k < 4; ++k) out[4 * s + k] = z[s][k];
    });
}

int orc_sensitivity(void* h, const double* e, const double* lam, double* g) {
    Oracle* o = static_cast<Oracle*>(h);
    return guard([&] {
        ComplexVec a(cptr(e), cptr(e) + o->ne), b(cptr(lam), cptr(lam) + o->ne);
        RealVec gv;
        o->sensitivity(a, b, gv);
        std::copy(gv.begin(), gv.end(), g);
    });
}

}  // extern "C"
