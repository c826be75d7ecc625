// Host mesh in the reference numbering (mesh.h:60-263): generators, CSR
// connectivity, geometry and face zones. The device layout is derived from
// it by the SFC renumbering in layout.cpp.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace mlr {

struct FaceZone {
    std::string name;
    bool interior = false;
    std::vector<uint32_t> faces;  // zone.h:57-102, generation order (interior sorted)
};

struct Mesh {
    MeshType type = MeshType::CartesianTri;
    uint32_t nx = 0, ny = 0;
    uint32_t n_cells = 0, n_nodes = 0;
    uint32_t n_faces = 0;       // allocated count (mesh.cpp:562 over-allocates for triangles)
    uint32_t n_faces_real = 0;  // ids actually generated
    std::vector<double> node_coords, cell_coords, cell_volume, face_area, face_normals;
    std::vector<uint32_t> nodes_of_cell, offsets_nodes_of_cell;
    std::vector<uint32_t> faces_of_cell, offsets_faces_of_cell;
    std::vector<uint32_t> nodes_of_face, offsets_nodes_of_face;
    std::vector<int32_t> cells_of_face;  // [f][2], -1 marks a boundary side
    std::vector<FaceZone> zones;

    uint32_t n_nodes_of_cell(uint32_t c) const { return offsets_nodes_of_cell[c + 1] - offsets_nodes_of_cell[c]; }
    uint32_t n_faces_of_cell(uint32_t c) const { return offsets_faces_of_cell[c + 1] - offsets_faces_of_cell[c]; }
    uint32_t node_of_cell(uint32_t c, int l) const { return nodes_of_cell[offsets_nodes_of_cell[c] + l]; }
    uint32_t face_of_cell(uint32_t c, int l) const { return faces_of_cell[offsets_faces_of_cell[c] + l]; }
    uint32_t node_of_face(uint32_t f, int l) const { return nodes_of_face[offsets_nodes_of_face[f] + l]; }
    double nc(uint32_t n, int d) const { return node_coords[2 * (size_t)n + d]; }
    double cc(uint32_t c, int d) const { return cell_coords[2 * (size_t)c + d]; }
    int32_t cof(uint32_t f, int s) const { return cells_of_face[2 * (size_t)f + s]; }
    int zone_index(const std::string& name) const;
    uint32_t max_faces_of_cell() const;

    void compute_geometry();  // mesh.cpp:167-261, called after node coordinates are final
};

// Generators. The jitter extension moves interior nodes of the structured
// generators by jitter * (dx, dy) * (2u - 1), u from splitmix64 keyed by
// (seed, node id); jitter = 0 reproduces the reference generator exactly.
Mesh make_cartesian(uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter = 0.0,
                    uint64_t seed = 0);
Mesh make_cartesian_tri(uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter = 0.0,
                        uint64_t seed = 0);
Mesh make_wedge(uint32_t nx, uint32_t ny, double Lx, double Ly);

// Build from caller arrays in the reference numbering (C-ABI path): faces
// with cells_of_face[f][1] == -1 must be listed in exactly one boundary zone.
Mesh make_from_arrays(uint32_t n_nodes, const double* node_coords, uint32_t n_cells,
                      const uint32_t* offsets_nodes_of_cell, const uint32_t* nodes_of_cell,
                      const uint32_t* faces_of_cell, uint32_t n_faces,
                      const uint32_t* nodes_of_face, const int32_t* cells_of_face,
                      uint32_t n_zones, const char* const* zone_names,
                      const uint32_t* zone_offsets, const uint32_t* zone_faces);

}  // namespace mlr
