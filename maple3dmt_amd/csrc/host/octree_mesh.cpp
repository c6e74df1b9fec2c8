// p4est-free octree builder; see octree_mesh.h for the reference mapping.
#include "octree_mesh.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <unordered_map>

namespace m3d {

namespace {

inline int32_t quad_len(int level) { return OctreeMesh::kRootLen >> level; }

// Physical box of an octant, evaluated with the same operation order as the
// reference callbacks (octree_mesh.cpp:139-160, 220-232, 285-297) so that
// predicate decisions on exact boundaries agree.
struct Box {
    double x0, y0, z0, x1, y1, z1;
};

Box octant_box(const RefinementParams& p, const Octant& q) {
    const double rl = static_cast<double>(OctreeMesh::kRootLen);
    const int32_t ql = quad_len(q.level);
    const double lx = p.domain_x_max - p.domain_x_min;
    const double ly = p.domain_y_max - p.domain_y_min;
    const double lz = p.domain_z_max - p.domain_z_min;
    Box b;
    b.x0 = p.domain_x_min + static_cast<double>(q.x) / rl * lx;
    b.y0 = p.domain_y_min + static_cast<double>(q.y) / rl * ly;
    b.z0 = p.domain_z_min + static_cast<double>(q.z) / rl * lz;
    b.x1 = p.domain_x_min + static_cast<double>(q.x + ql) / rl * lx;
    b.y1 = p.domain_y_min + static_cast<double>(q.y + ql) / rl * ly;
    b.z1 = p.domain_z_min + static_cast<double>(q.z + ql) / rl * lz;
    return b;
}

// Squared distance from a point to an axis-aligned box (0 inside),
// octree_mesh.cpp:115-127.
double box_point_dist2(double px, double py, double pz, const Box& b) {
    auto gap = [](double v, double lo, double hi) {
        return v < lo ? lo - v : (v > hi ? v - hi : 0.0);
    };
    const double dx = gap(px, b.x0, b.x1);
    const double dy = gap(py, b.y0, b.y1);
    const double dz = gap(pz, b.z0, b.z1);
    return dx * dx + dy * dy + dz * dz;
}

inline uint64_t octant_key(int level, int32_t x, int32_t y, int32_t z) {
    return (uint64_t(level) << 60) | (uint64_t(x) << 40) | (uint64_t(y) << 20) | uint64_t(z);
}

struct SkinLayer {
    double z_top, z_bot;
    int level;
};

}  // namespace

// One non-recursive refinement pass (p8est_refine with recursive=0): every
// leaf whose predicate holds is replaced by its 8 children in Morton order;
// children are not re-tested in the same pass.
void OctreeMesh::refine_pass(const std::function<bool(const Octant&)>& pred, bool* changed) {
    std::vector<Octant> next;
    next.reserve(leaves_.size() * 2);
    bool any = false;
    for (const Octant& q : leaves_) {
        const bool want = pred(q);  // evaluated for every leaf, in order
        if (want && q.level < kMaxDepth - 1) {
            const int32_t h = quad_len(q.level + 1);
            for (int child = 0; child < 8; ++child) {
                Octant c;
                c.x = q.x + ((child & 1) ? h : 0);
                c.y = q.y + ((child & 2) ? h : 0);
                c.z = q.z + ((child & 4) ? h : 0);
                c.level = static_cast<int8_t>(q.level + 1);
                next.push_back(c);
            }
            any = true;
        } else {
            next.push_back(q);
        }
    }
    leaves_.swap(next);
    if (changed) *changed = any;
}

// 2:1 balance across faces only (p8est_balance with P8EST_CONNECT_FACE,
// octree_mesh.cpp:508).  The fixpoint of "refine every leaf that has a face
// neighbour more than one level finer" is the unique coarsest face-balanced
// refinement, which is what p8est_balance produces.
void OctreeMesh::balance_face() {
    for (;;) {
        std::unordered_map<uint64_t, int> index;
        index.reserve(leaves_.size() * 2);
        for (int i = 0; i < static_cast<int>(leaves_.size()); ++i) {
            const Octant& q = leaves_[i];
            index.emplace(octant_key(q.level, q.x, q.y, q.z), i);
        }
        std::vector<char> mark(leaves_.size(), 0);
        bool any = false;
        for (const Octant& q : leaves_) {
            if (q.level < 2) continue;
            const int32_t len = quad_len(q.level);
            const int32_t probes[6][3] = {
                {q.x + len, q.y, q.z}, {q.x - len, q.y, q.z},
                {q.x, q.y + len, q.z}, {q.x, q.y - len, q.z},
                {q.x, q.y, q.z + len}, {q.x, q.y, q.z - len}};
            for (const auto& p : probes) {
                if (p[0] < 0 || p[1] < 0 || p[2] < 0 ||
                    p[0] >= kRootLen || p[1] >= kRootLen || p[2] >= kRootLen)
                    continue;
                for (int lv = 0; lv <= q.level - 2; ++lv) {
                    const int32_t m = ~(quad_len(lv) - 1);
                    auto it = index.find(octant_key(lv, p[0] & m, p[1] & m, p[2] & m));
                    if (it != index.end()) {
                        mark[it->second] = 1;
                        any = true;
                        break;
                    }
                }
            }
        }
        if (!any) break;
        int pos = 0;
        refine_pass([&](const Octant&) { return mark[pos++] != 0; }, nullptr);
    }
}

void OctreeMesh::sort_morton() {
    // Refinement replaces a leaf by its children in place, so the leaf list
    // stays in Morton order; this sort only guards that invariant.
    auto morton = [](const Octant& q) {
        uint64_t k = 0;
        for (int b = kMaxDepth - 1; b >= 0; --b) {
            k = (k << 3) | (uint64_t((q.z >> b) & 1) << 2) |
                (uint64_t((q.y >> b) & 1) << 1) | uint64_t((q.x >> b) & 1);
        }
        return k;
    };
    if (!std::is_sorted(leaves_.begin(), leaves_.end(),
                        [&](const Octant& a, const Octant& b) { return morton(a) < morton(b); }))
        std::sort(leaves_.begin(), leaves_.end(),
                  [&](const Octant& a, const Octant& b) { return morton(a) < morton(b); });
}

// Cell centres and sizes, octree_mesh.cpp:523-566.
void OctreeMesh::extract_cells() {
    const int n = num_cells();
    const double dx = params_.domain_x_max - params_.domain_x_min;
    const double dy = params_.domain_y_max - params_.domain_y_min;
    const double dz = params_.domain_z_max - params_.domain_z_min;
    const double rl = static_cast<double>(kRootLen);
    cx_.resize(n); cy_.resize(n); cz_.resize(n); csize_.resize(n);
    ctype_.resize(n, CellType::EARTH);
    for (int i = 0; i < n; ++i) {
        const Octant& q = leaves_[i];
        const int32_t ql = quad_len(q.level);
        const double cxu = (q.x + ql / 2.0) / rl;
        const double cyu = (q.y + ql / 2.0) / rl;
        const double czu = (q.z + ql / 2.0) / rl;
        const double hu = static_cast<double>(ql) / rl;
        cx_[i] = params_.domain_x_min + cxu * dx;
        cy_[i] = params_.domain_y_min + cyu * dy;
        cz_[i] = params_.domain_z_min + czu * dz;
        csize_[i] = hu * std::min({dx, dy, dz});
    }
}

void OctreeMesh::cell_size_xyz(int c, double& dx, double& dy, double& dz) const {
    const double hu = 1.0 / (1 << leaves_[c].level);
    dx = hu * (params_.domain_x_max - params_.domain_x_min);
    dy = hu * (params_.domain_y_max - params_.domain_y_min);
    dz = hu * (params_.domain_z_max - params_.domain_z_min);
}

void OctreeMesh::setup(const RefinementParams& params,
                       const std::vector<std::array<double, 3>>& station_xyz,
                       const std::vector<double>& frequencies) {
    params_ = params;
    if (params_.max_level > 17)
        throw std::invalid_argument("max_level above 17 is not supported");
    if (params_.min_level < 0 || params_.min_level > params_.max_level)
        throw std::invalid_argument("invalid min/max level");

    // Uniform base forest at min_level in Morton order (p8est_new_ext, :40-42).
    leaves_.clear();
    {
        const int lv = params_.min_level;
        const int32_t len = quad_len(lv);
        const uint64_t count = uint64_t(1) << (3 * lv);
        leaves_.reserve(count);
        for (uint64_t m = 0; m < count; ++m) {
            int32_t x = 0, y = 0, z = 0;
            for (int b = 0; b < lv; ++b) {
                x |= int32_t((m >> (3 * b)) & 1) << b;
                y |= int32_t((m >> (3 * b + 1)) & 1) << b;
                z |= int32_t((m >> (3 * b + 2)) & 1) << b;
            }
            leaves_.push_back(Octant{x * len, y * len, z * len, static_cast<int8_t>(lv)});
        }
    }
    const RefinementParams& p = params_;
    const double lx = p.domain_x_max - p.domain_x_min;
    const double ly = p.domain_y_max - p.domain_y_min;
    const double lz = p.domain_z_max - p.domain_z_min;

    // Station refinement passes (:129-200).
    if (!station_xyz.empty()) {
        const double diag = std::max({lx, ly, lz});
        const double h_finest = diag / (1 << p.station_refine_level);
        const double air_max_z = 2.0 * h_finest;
        const double r2 = p.station_refine_radius * p.station_refine_radius;
        auto pred = [&](const Octant& q) {
            if (q.level >= p.station_refine_level) return false;
            const Box b = octant_box(p, q);
            for (const auto& st : station_xyz) {
                if (b.z0 > st[2] + air_max_z) continue;
                if (box_point_dist2(st[0], st[1], st[2], b) < r2) return true;
            }
            return false;
        };
        for (int pass = 0; pass < p.max_level - p.min_level; ++pass) refine_pass(pred, nullptr);
    }

    // Region refinement passes (:211-261).
    for (const RefineRegion& reg : p.refine_regions) {
        auto pred = [&](const Octant& q) {
            if (q.level >= reg.level) return false;
            const Box b = octant_box(p, q);
            const double pad = reg.padding;
            if (b.x1 < reg.x_min - pad || b.x0 > reg.x_max + pad) return false;
            if (b.y1 < reg.y_min - pad || b.y0 > reg.y_max + pad) return false;
            if (b.z1 < reg.z_min - pad || b.z0 > reg.z_max + pad) return false;
            return true;
        };
        for (int pass = 0; pass < reg.level - p.min_level; ++pass) refine_pass(pred, nullptr);
    }

    // Skin-depth layer refinement (:316-493).
    if (!frequencies.empty()) {
        const double f_min = *std::min_element(frequencies.begin(), frequencies.end());
        const double delta_max = std::sqrt(2.0 / (kTwoPi * f_min * kMu0 * p.sigma_bg));
        double rx0, rx1, ry0, ry1;
        const double pad_h = 0.5 * delta_max + p.transition_padding;
        if (p.station_x_max > p.station_x_min) {
            rx0 = p.station_x_min - pad_h; rx1 = p.station_x_max + pad_h;
            ry0 = p.station_y_min - pad_h; ry1 = p.station_y_max + pad_h;
        } else {
            const double hx = (p.domain_x_max - p.domain_x_min) * 0.2;
            const double hy = (p.domain_y_max - p.domain_y_min) * 0.2;
            rx0 = -hx - pad_h; rx1 = hx + pad_h;
            ry0 = -hy - pad_h; ry1 = hy + pad_h;
        }
        rx0 = std::max(rx0, p.domain_x_min); rx1 = std::min(rx1, p.domain_x_max);
        ry0 = std::max(ry0, p.domain_y_min); ry1 = std::min(ry1, p.domain_y_max);

        const int off = (p.max_level >= 9) ? 2 : 1;
        const int L1 = std::max(p.min_level + 1, p.max_level - off);
        const int L2 = std::max(p.min_level + 1, p.max_level - off - 1);
        const int L3 = std::max(p.min_level, p.max_level - off - 2);
        double z_int = -3.0 * delta_max;
        if (p.max_interest_depth > 0) z_int = -p.max_interest_depth;

        std::vector<SkinLayer> layers;
        const double z1b = std::max(-0.5 * delta_max, z_int);
        if (z1b < 0) layers.push_back({0.0, z1b, L1});
        const double z2t = -0.5 * delta_max, z2b = std::max(-1.5 * delta_max, z_int);
        if (z2t > z_int && z2b < z2t) layers.push_back({z2t, z2b, L2});
        const double z3t = -1.5 * delta_max, z3b = std::max(-3.0 * delta_max, z_int);
        if (z3t > z_int && z3b < z3t && L3 > p.min_level) layers.push_back({z3t, z3b, L3});
        const int L_deep = p.min_level + 1;
        if (z_int > p.domain_z_min && L_deep > p.min_level)
            layers.push_back({z_int, p.domain_z_min, L_deep});
        const int La1 = std::max(p.min_level + 1, L1 - 1);
        const int La2 = std::max(p.min_level, La1 - 1);
        const int La3 = std::max(p.min_level, La2 - 1);
        const double za1 = p.air_surface_thickness;
        const double za2 = za1 + p.air_mid_thickness;
        const double za3 = za2 + p.air_upper_thickness;
        layers.push_back({za1, 0.0, La1});
        layers.push_back({za2, za1, La2});
        layers.push_back({za3, za2, La3});

        auto pred = [&](const Octant& q) {
            const Box b = octant_box(p, q);
            if (b.x1 < rx0 || b.x0 > rx1) return false;
            if (b.y1 < ry0 || b.y0 > ry1) return false;
            int target = p.min_level;
            for (const SkinLayer& l : layers)
                if (b.z1 > l.z_bot && b.z0 < l.z_top) target = std::max(target, l.level);
            return q.level < target;
        };
        for (int pass = 0; pass < p.max_level - p.min_level; ++pass) {
            bool changed = false;
            refine_pass(pred, &changed);
            if (!changed) break;
        }
    }

    balance_face();
    sort_morton();
    extract_cells();
}

void OctreeMesh::set_terrain(std::function<double(double, double)> dem) {
    terrain_ = std::move(dem);
    // Terrain refinement is a no-op in the reference; re-balancing an already
    // balanced forest changes nothing.
    balance_face();
    extract_cells();
    classify();
}

// Earth / air / ocean by terrain height, octree_mesh.cpp:656-694.
void OctreeMesh::classify() {
    const int n = num_cells();
    if (!terrain_) {
        for (int i = 0; i < n; ++i) ctype_[i] = (cz_[i] < 0.0) ? CellType::EARTH : CellType::AIR;
        return;
    }
    for (int i = 0; i < n; ++i) {
        const double tz = terrain_(cx_[i], cy_[i]);
        double ddx, ddy, h;
        cell_size_xyz(i, ddx, ddy, h);
        if (cz_[i] + h / 2.0 < tz) {
            ctype_[i] = CellType::EARTH;
        } else if (cz_[i] - h / 2.0 > tz) {
            ctype_[i] = CellType::AIR;
        } else {
            double frac = (tz - (cz_[i] - h / 2.0)) / h;
            frac = std::clamp(frac, 0.0, 1.0);
            ctype_[i] = (frac > 0.5) ? CellType::EARTH : CellType::AIR;
        }
        if (tz < 0.0 && ctype_[i] == CellType::AIR && cz_[i] < 0.0) ctype_[i] = CellType::OCEAN;
    }
}

}  // namespace m3d
