// Staggered-grid enumeration; see staggered_grid.h for the reference mapping.
#include "staggered_grid.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "flat_map.h"

namespace m3d {

namespace {

struct Quantizer {
    double dx, dy, dz, x0, y0, z0, quant;
    // (quantised midpoint, direction) packed in 64 bits: 20 bits per axis,
    // 2 bits of direction / normal.  max_level <= 17 keeps every coordinate
    // below 2^20.
    uint64_t key(double px, double py, double pz, int dir) const {
        const int64_t ix = static_cast<int64_t>(std::round((px - x0) / dx * quant));
        const int64_t iy = static_cast<int64_t>(std::round((py - y0) / dy * quant));
        const int64_t iz = static_cast<int64_t>(std::round((pz - z0) / dz * quant));
        return (uint64_t(ix & 0xFFFFF) << 42) | (uint64_t(iy & 0xFFFFF) << 22) |
               (uint64_t(iz & 0xFFFFF) << 2) | uint64_t(dir & 3);
    }
};

Quantizer make_quantizer(const OctreeMesh& mesh) {
    const auto& p = mesh.params();
    Quantizer q;
    q.dx = p.domain_x_max - p.domain_x_min;
    q.dy = p.domain_y_max - p.domain_y_min;
    q.dz = p.domain_z_max - p.domain_z_min;
    q.x0 = p.domain_x_min;
    q.y0 = p.domain_y_min;
    q.z0 = p.domain_z_min;
    q.quant = static_cast<double>(1 << (p.max_level + 2));
    return q;
}

struct Half {
    double cx, cy, cz, hx, hy, hz;
};

Half cell_half(const OctreeMesh& mesh, const Quantizer& q, int c) {
    Half h;
    mesh.cell_center(c, h.cx, h.cy, h.cz);
    const double hu = 1.0 / (1 << mesh.cell_level(c));
    h.hx = hu * q.dx / 2.0;
    h.hy = hu * q.dy / 2.0;
    h.hz = hu * q.dz / 2.0;
    return h;
}

}  // namespace

void StaggeredGrid::build(const OctreeMesh& mesh) {
    num_cells_ = mesh.num_cells();
    const Quantizer qz = make_quantizer(mesh);

    edge_x.clear(); edge_y.clear(); edge_z.clear(); edge_len.clear();
    edge_dir.clear(); edge_adj.clear();
    FlatMap<uint64_t> edge_map(size_t(num_cells_) * 4);

    auto add_edge = [&](double x, double y, double z, int dir, double len, int c) {
        const auto r = edge_map.insert(qz.key(x, y, z, dir), num_edges());
        const int id = r.first;
        if (r.second) {
            edge_x.push_back(x); edge_y.push_back(y); edge_z.push_back(z);
            edge_dir.push_back(static_cast<int8_t>(dir));
            edge_len.push_back(len);
            edge_adj.emplace_back();
            edge_adj.back().reserve(4);
        }
        edge_adj[id].push_back(c);
    };

    // Phase 1: 12 edges per cell in the fixed per-cell order of
    // staggered_grid.cpp:85-120.
    for (int c = 0; c < num_cells_; ++c) {
        const Half h = cell_half(mesh, qz, c);
        const double ys[4] = {h.cy - h.hy, h.cy - h.hy, h.cy + h.hy, h.cy + h.hy};
        const double zs[4] = {h.cz - h.hz, h.cz + h.hz, h.cz - h.hz, h.cz + h.hz};
        for (int i = 0; i < 4; ++i) add_edge(h.cx, ys[i], zs[i], 0, 2.0 * h.hx, c);
        const double xs[4] = {h.cx - h.hx, h.cx - h.hx, h.cx + h.hx, h.cx + h.hx};
        for (int i = 0; i < 4; ++i) add_edge(xs[i], h.cy, zs[i], 1, 2.0 * h.hy, c);
        const double xz[4] = {h.cx - h.hx, h.cx + h.hx, h.cx - h.hx, h.cx + h.hx};
        const double yz[4] = {h.cy - h.hy, h.cy - h.hy, h.cy + h.hy, h.cy + h.hy};
        for (int i = 0; i < 4; ++i) add_edge(xz[i], yz[i], h.cz, 2, 2.0 * h.hz, c);
    }

    // Phase 2: 6 faces per cell (+x,-x,+y,-y,+z,-z) with their 4 bounding
    // edges, staggered_grid.cpp:128-224.
    face_x.clear(); face_y.clear(); face_z.clear(); face_area.clear();
    face_normal.clear(); face_minus.clear(); face_plus.clear(); face_edges.clear();
    FlatMap<uint64_t> face_map(size_t(num_cells_) * 4);

    struct ERef { double x, y, z; int dir; };
    for (int c = 0; c < num_cells_; ++c) {
        const Half h = cell_half(mesh, qz, c);
        const double cx = h.cx, cy = h.cy, cz = h.cz, hx = h.hx, hy = h.hy, hz = h.hz;
        struct FDef { double x, y, z; int normal; double area; ERef e[4]; };
        const FDef fd[6] = {
            {cx + hx, cy, cz, 0, 4 * hy * hz,
             {{cx + hx, cy, cz - hz, 1}, {cx + hx, cy, cz + hz, 1},
              {cx + hx, cy - hy, cz, 2}, {cx + hx, cy + hy, cz, 2}}},
            {cx - hx, cy, cz, 0, 4 * hy * hz,
             {{cx - hx, cy, cz - hz, 1}, {cx - hx, cy, cz + hz, 1},
              {cx - hx, cy - hy, cz, 2}, {cx - hx, cy + hy, cz, 2}}},
            {cx, cy + hy, cz, 1, 4 * hx * hz,
             {{cx, cy + hy, cz - hz, 0}, {cx, cy + hy, cz + hz, 0},
              {cx - hx, cy + hy, cz, 2}, {cx + hx, cy + hy, cz, 2}}},
            {cx, cy - hy, cz, 1, 4 * hx * hz,
             {{cx, cy - hy, cz - hz, 0}, {cx, cy - hy, cz + hz, 0},
              {cx - hx, cy - hy, cz, 2}, {cx + hx, cy - hy, cz, 2}}},
            {cx, cy, cz + hz, 2, 4 * hx * hy,
             {{cx, cy - hy, cz + hz, 0}, {cx, cy + hy, cz + hz, 0},
              {cx - hx, cy, cz + hz, 1}, {cx + hx, cy, cz + hz, 1}}},
            {cx, cy, cz - hz, 2, 4 * hx * hy,
             {{cx, cy - hy, cz - hz, 0}, {cx, cy + hy, cz - hz, 0},
              {cx - hx, cy, cz - hz, 1}, {cx + hx, cy, cz - hz, 1}}},
        };
        for (int f = 0; f < 6; ++f) {
            const FDef& d = fd[f];
            const auto r = face_map.insert(qz.key(d.x, d.y, d.z, d.normal), num_faces());
            const int fid = r.first;
            if (r.second) {
                std::vector<int> eids;
                eids.reserve(4);
                for (int e = 0; e < 4; ++e) {
                    const int ei = edge_map.find(qz.key(d.e[e].x, d.e[e].y, d.e[e].z, d.e[e].dir));
                    if (ei >= 0) eids.push_back(ei);
                }
                face_x.push_back(d.x); face_y.push_back(d.y); face_z.push_back(d.z);
                face_normal.push_back(static_cast<int8_t>(d.normal));
                face_area.push_back(d.area);
                face_minus.push_back(-1); face_plus.push_back(-1);
                face_edges.push_back(std::move(eids));
            }
            if (f == 0 || f == 2 || f == 4) face_minus[fid] = c;
            else face_plus[fid] = c;
        }
    }

    detect_hanging(mesh);
    build_cell_edge_map();
}

// Coarse/fine interface detection, staggered_grid.cpp:241-391.
void StaggeredGrid::detect_hanging(const OctreeMesh& mesh) {
    hanging_faces.clear();
    hanging_edges.clear();
    const auto& par = mesh.params();
    if (par.min_level >= par.max_level) return;

    const Quantizer qz = make_quantizer(mesh);
    const double min_h = std::min({qz.dx / (1 << par.max_level), qz.dy / (1 << par.max_level),
                                   qz.dz / (1 << par.max_level)});
    const double btol = min_h * 0.01;

    // Faces and edges have distinct keys (they were enumerated by them).
    FlatMap<uint64_t> face_lk(static_cast<size_t>(num_faces()));
    FlatMap<uint64_t> edge_lk(static_cast<size_t>(num_edges()));
    for (int f = 0; f < num_faces(); ++f) face_lk.insert(qz.key(face_x[f], face_y[f], face_z[f], face_normal[f]), f);
    for (int e = 0; e < num_edges(); ++e) edge_lk.insert(qz.key(edge_x[e], edge_y[e], edge_z[e], edge_dir[e]), e);

    std::vector<char> face_done(num_faces(), 0);
    std::vector<char> edge_hanging(num_edges(), 0);

    for (int f = 0; f < num_faces(); ++f) {
        if (face_done[f]) continue;
        if (face_minus[f] >= 0 && face_plus[f] >= 0) continue;
        if (face_minus[f] < 0 && face_plus[f] < 0) continue;
        const int n = face_normal[f];
        bool on_bnd = false;
        if (n == 0) on_bnd = face_x[f] < par.domain_x_min + btol || face_x[f] > par.domain_x_max - btol;
        else if (n == 1) on_bnd = face_y[f] < par.domain_y_min + btol || face_y[f] > par.domain_y_max - btol;
        else on_bnd = face_z[f] < par.domain_z_min + btol || face_z[f] > par.domain_z_max - btol;
        if (on_bnd) continue;

        const int ccell = (face_minus[f] >= 0) ? face_minus[f] : face_plus[f];
        const Half ch = cell_half(mesh, qz, ccell);
        const double fhx = ch.hx / 2, fhy = ch.hy / 2, fhz = ch.hz / 2;

        std::array<int, 4> subs{};
        bool found = true;
        int idx = 0;
        for (int s1 = -1; s1 <= 1 && found; s1 += 2) {
            for (int s2 = -1; s2 <= 1; s2 += 2) {
                double sx = face_x[f], sy = face_y[f], sz = face_z[f];
                if (n == 0) { sy += s1 * fhy; sz += s2 * fhz; }
                else if (n == 1) { sx += s1 * fhx; sz += s2 * fhz; }
                else { sx += s1 * fhx; sy += s2 * fhy; }
                const int ff = face_lk.find(qz.key(sx, sy, sz, n));
                if (ff < 0 || face_done[ff]) { found = false; break; }
                subs[idx++] = ff;
            }
        }
        if (!found) continue;
        for (int ff : subs)
            if (face_minus[ff] >= 0 && face_plus[ff] >= 0) { found = false; break; }
        if (!found) continue;

        hanging_faces.push_back({f, subs});
        face_done[f] = 1;
        for (int ff : subs) face_done[ff] = 1;
        for (int ff : subs) {
            if (face_minus[ff] < 0) face_minus[ff] = ccell;
            else if (face_plus[ff] < 0) face_plus[ff] = ccell;
        }
        face_minus[f] = -1;
        face_plus[f] = -1;
        for (int ff : subs) {
            for (int eid : face_edges[ff]) {
                if (eid < 0) continue;
                auto& adj = edge_adj[eid];
                if (std::find(adj.begin(), adj.end(), ccell) == adj.end()) adj.push_back(ccell);
            }
        }
        for (int eid : face_edges[f]) {
            if (eid < 0 || edge_hanging[eid]) continue;
            const int d = edge_dir[eid];
            const double ql = edge_len[eid] / 4.0;
            double a[3] = {edge_x[eid], edge_y[eid], edge_z[eid]};
            double b[3] = {edge_x[eid], edge_y[eid], edge_z[eid]};
            a[d] -= ql;
            b[d] += ql;
            const int i1 = edge_lk.find(qz.key(a[0], a[1], a[2], d));
            const int i2 = edge_lk.find(qz.key(b[0], b[1], b[2], d));
            if (i1 >= 0 && i2 >= 0) {
                hanging_edges.push_back({eid, {i1, i2}});
                edge_hanging[eid] = 1;
            }
        }
    }
}

void StaggeredGrid::build_cell_edge_map() {
    cell_edges.assign(num_cells_, {});
    for (int e = 0; e < num_edges(); ++e) {
        const auto& adj = edge_adj[e];
        if (adj.empty()) continue;
        const double frac = 1.0 / static_cast<double>(adj.size());
        for (int c : adj)
            if (c >= 0 && c < num_cells_) cell_edges[c].push_back({e, frac});
    }
}

}  // namespace m3d
