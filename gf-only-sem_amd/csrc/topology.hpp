// topology.hpp -- general hexahedral meshes: GMSH reader, the reference's
// /topology group (vertices, edges, surfaces, cells with signed 1-based ids)
// and each cell's corners in GMSH order.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace gfx::topo {

// Points and 8-node hexahedra (0-based point indices, GMSH corner order).
struct HexMesh {
    std::vector<double> xyz;   // [n_point][3]
    std::vector<int64_t> hex;  // [n_cell][8]
};

struct Topology {
    std::vector<double> vertex_to_coord;   // [n_vertex][3]
    std::vector<int64_t> edge_to_vertex;   // [n_edge][2], 1-based, low < high
    std::vector<int64_t> surface_to_edge;  // [n_surface][4], signed 1-based edge ids
    std::vector<int64_t> cell_to_surface;  // [n_cell][6], signed 1-based; -z +z -y +y -x +x
    int64_t n_vertex = 0, n_edge = 0, n_surface = 0, n_cell = 0;
};

HexMesh read_gmsh(const std::string& path);
Topology extract(const HexMesh& m);
void write_topology(const std::string& path, const Topology& t);
void write_auxiliary(const std::string& path, const Topology& t);
Topology read_topology(const std::string& path);
// 1-based vertex ids of cell c in GMSH corner order.
std::array<int64_t, 8> cell_vertices(const Topology& t, int64_t c);

}  // namespace gfx::topo
