// Octree mesh for the octree finite-volume MT solve, built without p4est.
//
// A single unit-cube tree mapped onto the physical domain, refined by
// non-recursive one-level passes, 2:1 face-balanced, and stored in Morton
// (z-major, then y, then x) leaf order.  This reproduces the behaviour of
// OctreeMesh::setup / set_terrain in src/octree/octree_mesh.cpp:25-78 with the
// refinement predicates of :129-183 (stations), :211-241 (regions) and
// :279-313 with the layer table of :316-493 (skin depth), the face balance of
// :507-518 and the flat-terrain classification of :656-709.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <vector>

namespace m3d {

constexpr double kMu0 = 4.0e-7 * 3.14159265358979323846;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;

// Mirrors RefineRegion (include/maple3dmt/octree/octree_mesh.h:31-35).
struct RefineRegion {
    double x_min, x_max, y_min, y_max, z_min, z_max;
    int level;
    double padding;
};

// Mirrors RefinementParams (include/maple3dmt/octree/octree_mesh.h:37-90),
// same field names and defaults.
struct RefinementParams {
    double domain_x_min = 0, domain_x_max = 0;
    double domain_y_min = 0, domain_y_max = 0;
    double domain_z_min = 0, domain_z_max = 0;
    int min_level = 3;
    int max_level = 8;
    double station_refine_radius = 5000.0;
    int station_refine_level = 7;
    std::vector<RefineRegion> refine_regions;
    int terrain_refine_level = 6;
    int skin_depth_levels = 2;
    double sigma_bg = 0.01;
    double max_interest_depth = 0.0;
    double air_surface_thickness = 2000.0;
    double air_mid_thickness = 8000.0;
    double air_upper_thickness = 20000.0;
    double transition_padding = 0.0;
    double station_x_min = 0, station_x_max = 0;
    double station_y_min = 0, station_y_max = 0;
};

enum class CellType : int { EARTH = 0, AIR = 1, OCEAN = 2 };

// Leaf octant in integer coordinates of a 2^19 root (p8est's P8EST_ROOT_LEN).
struct Octant {
    int32_t x, y, z;
    int8_t level;
};

class OctreeMesh {
public:
    static constexpr int kMaxDepth = 19;
    static constexpr int32_t kRootLen = int32_t(1) << kMaxDepth;

    // Build: uniform base at min_level, station / region / skin-depth passes,
    // face balance, leaf extraction (octree_mesh.cpp:25-63).
    void setup(const RefinementParams& params,
               const std::vector<std::array<double, 3>>& station_xyz,
               const std::vector<double>& frequencies);

    // Terrain classification (octree_mesh.cpp:65-78, 656-709).  The
    // reference's terrain refinement is a no-op (:498-502).
    void set_terrain(std::function<double(double, double)> dem);

    int num_cells() const { return static_cast<int>(leaves_.size()); }
    const RefinementParams& params() const { return params_; }
    const std::vector<Octant>& leaves() const { return leaves_; }

    void cell_center(int c, double& x, double& y, double& z) const {
        x = cx_[c]; y = cy_[c]; z = cz_[c];
    }
    double cell_size(int c) const { return csize_[c]; }
    void cell_size_xyz(int c, double& dx, double& dy, double& dz) const;
    int cell_level(int c) const { return leaves_[c].level; }
    CellType cell_type(int c) const { return ctype_[c]; }
    void set_cell_type(int c, CellType t) { ctype_[c] = t; }

private:
    RefinementParams params_;
    std::vector<Octant> leaves_;
    std::vector<double> cx_, cy_, cz_, csize_;
    std::vector<CellType> ctype_;
    std::function<double(double, double)> terrain_;

    void refine_pass(const std::function<bool(const Octant&)>& pred, bool* changed);
    void balance_face();
    void sort_morton();
    void extract_cells();
    void classify();
};

}  // namespace m3d
