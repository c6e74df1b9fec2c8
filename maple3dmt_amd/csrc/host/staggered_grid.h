// Yee staggered grid on the octree: edge / face enumeration, hanging-face and
// hanging-edge detection, cell -> edge map.
//
// Follows StaggeredGrid::build (src/octree/staggered_grid.cpp:70-236),
// detect_hanging_ (:241-391) and build_cell_edge_map_ (:393-409).  Edges and
// faces are deduplicated by (quantised midpoint, direction) keys with
// quant = 2^(max_level+2) (:48) and numbered by first appearance in cell
// order, which fixes the DOF numbering and therefore the natural SSOR order.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "octree_mesh.h"

namespace m3d {

struct HangingFace {
    int coarse_face;
    std::array<int, 4> fine_faces;
};

struct HangingEdge {
    int coarse_edge;
    std::array<int, 2> fine_edges;
};

struct CellEdge {
    int edge_id;
    double volume_fraction;
};

class StaggeredGrid {
public:
    void build(const OctreeMesh& mesh);

    int num_edges() const { return static_cast<int>(edge_dir.size()); }
    int num_faces() const { return static_cast<int>(face_normal.size()); }
    int num_cells() const { return num_cells_; }

    // Edges: midpoint, direction (0=x,1=y,2=z), length, adjacent cells.
    std::vector<double> edge_x, edge_y, edge_z, edge_len;
    std::vector<int8_t> edge_dir;
    std::vector<std::vector<int>> edge_adj;

    // Faces: centre, normal, area, minus / plus cell (-1 = none), bounding
    // edges (up to 4, those that exist).
    std::vector<double> face_x, face_y, face_z, face_area;
    std::vector<int8_t> face_normal;
    std::vector<int> face_minus, face_plus;
    std::vector<std::vector<int>> face_edges;

    std::vector<HangingFace> hanging_faces;
    std::vector<HangingEdge> hanging_edges;
    std::vector<std::vector<CellEdge>> cell_edges;

private:
    int num_cells_ = 0;
    void detect_hanging(const OctreeMesh& mesh);
    void build_cell_edge_map();
};

}  // namespace m3d
