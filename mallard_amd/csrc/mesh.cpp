#include "mesh.hpp"

#include <algorithm>
#include <cmath>

namespace mlr {

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Regular (nx+1) x (ny+1) node lattice, node id i*(ny+1)+j
// (mesh.cpp:341-351, :595-605) with the optional interior-node jitter.
void lattice(Mesh& m, uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter, uint64_t seed) {
    const double dx = Lx / nx, dy = Ly / ny;
    m.node_coords.assign(2 * (size_t)m.n_nodes, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nx + 1; ++i)
        for (uint32_t j = 0; j < ny + 1; ++j) {
            const size_t n = (size_t)i * (ny + 1) + j;
            m.node_coords[2 * n] = (uint32_t)i * dx;
            m.node_coords[2 * n + 1] = j * dy;
        }
    if (jitter == 0.0) return;
    const uint64_t key = splitmix64(seed);
#pragma omp parallel for schedule(static)
    for (int64_t i = 1; i < (int64_t)nx; ++i)
        for (uint32_t j = 1; j < ny; ++j) {
            const uint64_t id = (uint64_t)i * (ny + 1) + j;
            const double u1 = (double)(splitmix64(key + 2 * id) >> 11) * 0x1.0p-53;
            const double u2 = (double)(splitmix64(key + 2 * id + 1) >> 11) * 0x1.0p-53;
            m.node_coords[2 * id] += jitter * dx * (2.0 * u1 - 1.0);
            m.node_coords[2 * id + 1] += jitter * dy * (2.0 * u2 - 1.0);
        }
}

void fixed_csr(Mesh& m, uint32_t nodes_per_cell) {
    m.nodes_of_cell.assign((size_t)m.n_cells * nodes_per_cell, 0);
    m.faces_of_cell.assign((size_t)m.n_cells * nodes_per_cell, 0);
    m.offsets_nodes_of_cell.resize(m.n_cells + 1);
    m.offsets_faces_of_cell.resize(m.n_cells + 1);
    for (uint32_t c = 0; c <= m.n_cells; ++c) {
        m.offsets_nodes_of_cell[c] = c * nodes_per_cell;
        m.offsets_faces_of_cell[c] = c * nodes_per_cell;
    }
    // every allocated face (phantom ones too) has two node slots, zero-filled
    m.nodes_of_face.assign(2 * (size_t)m.n_faces, 0);
    m.offsets_nodes_of_face.resize(m.n_faces + 1);
    for (uint32_t f = 0; f <= m.n_faces; ++f) m.offsets_nodes_of_face[f] = 2 * f;
    m.cells_of_face.assign(2 * (size_t)m.n_faces, 0);
}

void zones_from(Mesh& m, std::vector<uint8_t>& is_interior, std::vector<uint32_t> fr,
                std::vector<uint32_t> ft, std::vector<uint32_t> fl, std::vector<uint32_t> fb) {
    // interior zone: ascending unique face ids (mesh.cpp:731-733)
    FaceZone zi{"interior", true, {}};
    for (uint32_t f = 0; f < (uint32_t)is_interior.size(); ++f)
        if (is_interior[f]) zi.faces.push_back(f);
    m.zones.clear();
    m.zones.push_back(std::move(zi));
    m.zones.push_back({"right", false, std::move(fr)});
    m.zones.push_back({"top", false, std::move(ft)});
    m.zones.push_back({"left", false, std::move(fl)});
    m.zones.push_back({"bottom", false, std::move(fb)});
}

}  // namespace

int Mesh::zone_index(const std::string& name) const {
    for (size_t z = 0; z < zones.size(); ++z)
        if (zones[z].name == name) return (int)z;
    return -1;
}

uint32_t Mesh::max_faces_of_cell() const {
    uint32_t k = 0;
    for (uint32_t c = 0; c < n_cells; ++c) k = std::max(k, n_faces_of_cell(c));
    return k;
}

// mesh.cpp:218-261 and :167-216, same operation order.
void Mesh::compute_geometry() {
    face_area.assign(n_faces, 0.0);
    face_normals.assign(2 * (size_t)n_faces, 0.0);
    cell_volume.assign(n_cells, 0.0);
    cell_coords.assign(2 * (size_t)n_cells, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < (int64_t)n_faces; ++f) {
        const uint32_t a = node_of_face(f, 0), b = node_of_face(f, 1);
        const double ddx = nc(b, 0) - nc(a, 0), ddy = nc(b, 1) - nc(a, 1);
        face_area[f] = std::sqrt(ddx * ddx + ddy * ddy);
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)n_cells; ++c) {
        const uint32_t k = n_nodes_of_cell(c);
        auto tri = [&](uint32_t a, uint32_t b, uint32_t d) {
            return 0.5 * std::fabs(nc(a, 0) * (nc(b, 1) - nc(d, 1)) + nc(b, 0) * (nc(d, 1) - nc(a, 1)) +
                                   nc(d, 0) * (nc(a, 1) - nc(b, 1)));
        };
        if (k == 3) {
            cell_volume[c] = tri(node_of_cell(c, 0), node_of_cell(c, 1), node_of_cell(c, 2));
        } else if (k == 4) {
            const double a1 = tri(node_of_cell(c, 0), node_of_cell(c, 1), node_of_cell(c, 2));
            const double a2 = tri(node_of_cell(c, 0), node_of_cell(c, 2), node_of_cell(c, 3));
            cell_volume[c] = a1 + a2;
        } else {
            cell_volume[c] = NAN;
        }
        double x = 0.0, y = 0.0;
        for (uint32_t l = 0; l < k; ++l) {
            x += nc(node_of_cell(c, l), 0);
            y += nc(node_of_cell(c, l), 1);
        }
        cell_coords[2 * c] = x / k;
        cell_coords[2 * c + 1] = y / k;
    }
    for (uint32_t c = 0; c < n_cells; ++c) {
        const uint32_t k = n_nodes_of_cell(c);
        if (k != 3 && k != 4) fail("Unknown cell type.");
    }
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < (int64_t)n_faces; ++f) {
        const uint32_t a = node_of_face(f, 0), b = node_of_face(f, 1);
        const double x0 = nc(a, 0), y0 = nc(a, 1), x1 = nc(b, 0), y1 = nc(b, 1);
        const double ddx = x1 - x0, ddy = y1 - y0;
        const double mag = std::sqrt(ddx * ddx + ddy * ddy);
        double n0 = ddy / mag * face_area[f];
        double n1 = -ddx / mag * face_area[f];
        const int32_t c0 = cof(f, 0);
        const double d = (0.5 * (x0 + x1) - cc(c0, 0)) * n0 + (0.5 * (y0 + y1) - cc(c0, 1)) * n1;
        if (d < 0) { n0 *= -1; n1 *= -1; }
        face_normals[2 * f] = n0;
        face_normals[2 * f + 1] = n1;
    }
}

// mesh.cpp:305-557
Mesh make_cartesian(uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter, uint64_t seed) {
    if (nx == 0 || ny == 0) fail("mesh needs Nx, Ny > 0");
    Mesh m;
    m.type = MeshType::Cartesian;
    m.nx = nx; m.ny = ny;
    m.n_cells = nx * ny;
    m.n_nodes = (nx + 1) * (ny + 1);
    m.n_faces = 2 * m.n_cells + nx + ny;
    m.n_faces_real = m.n_faces;
    lattice(m, nx, ny, Lx, Ly, jitter, seed);
    fixed_csr(m, 4);
    std::vector<uint8_t> inner(m.n_faces, 0);
    std::vector<uint32_t> fr, ft, fl, fb;
    for (uint32_t c = 0; c < m.n_cells; ++c) {
        const uint32_t ic = c / ny, jc = c % ny;
        const uint32_t ntr = (ic + 1) * (ny + 1) + jc + 1, ntl = ic * (ny + 1) + jc + 1;
        const uint32_t nbl = ic * (ny + 1) + jc, nbr = (ic + 1) * (ny + 1) + jc;
        const uint32_t fR = (2 * ny + 1) * (ic + 1) + jc, fT = (2 * ny + 1) * ic + jc + ny + 1;
        const uint32_t fL = (2 * ny + 1) * ic + jc, fB = (2 * ny + 1) * ic + jc + ny;
        uint32_t* nn = &m.nodes_of_cell[4 * (size_t)c];
        nn[0] = ntr; nn[1] = ntl; nn[2] = nbl; nn[3] = nbr;
        uint32_t* ff = &m.faces_of_cell[4 * (size_t)c];
        ff[0] = fR; ff[1] = fT; ff[2] = fL; ff[3] = fB;
        auto set = [&](uint32_t f, int32_t a, int32_t b, uint32_t na, uint32_t nb) {
            m.cells_of_face[2 * (size_t)f] = a; m.cells_of_face[2 * (size_t)f + 1] = b;
            m.nodes_of_face[2 * (size_t)f] = na; m.nodes_of_face[2 * (size_t)f + 1] = nb;
        };
        set(fR, c, ic == nx - 1 ? -1 : (int32_t)(c + ny), nbr, ntr);
        set(fT, c, jc == ny - 1 ? -1 : (int32_t)(c + 1), ntr, ntl);
        set(fL, c, ic == 0 ? -1 : (int32_t)(c - ny), ntl, nbl);
        set(fB, c, jc == 0 ? -1 : (int32_t)(c - 1), nbl, nbr);
        if (ic == nx - 1) fr.push_back(fR); else inner[fR] = 1;
        if (jc == ny - 1) ft.push_back(fT); else inner[fT] = 1;
        if (ic == 0) fl.push_back(fL); else inner[fL] = 1;
        if (jc == 0) fb.push_back(fB); else inner[fB] = 1;
    }
    zones_from(m, inner, std::move(fr), std::move(ft), std::move(fl), std::move(fb));
    m.compute_geometry();
    return m;
}

// mesh.cpp:559-826. Triangle node order (br, tr, bl) / (tl, bl, tr) and
// face order (b, r, d) / (t, l, d) fix each cell's local frame.
Mesh make_cartesian_tri(uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter, uint64_t seed) {
    if (nx == 0 || ny == 0) fail("mesh needs Nx, Ny > 0");
    Mesh m;
    m.type = MeshType::CartesianTri;
    m.nx = nx; m.ny = ny;
    m.n_cells = 2 * nx * ny;
    m.n_nodes = (nx + 1) * (ny + 1);
    m.n_faces = 3 * m.n_cells + nx + ny;
    m.n_faces_real = 3 * nx * ny + nx + ny;
    lattice(m, nx, ny, Lx, Ly, jitter, seed);
    fixed_csr(m, 3);
    std::vector<uint8_t> inner(m.n_faces, 0);
    std::vector<uint32_t> fr, ft, fl, fb;
    for (uint32_t q = 0; q < nx * ny; ++q) {
        const uint32_t ic = q / ny, jc = q % ny;
        const uint32_t cr = 2 * q, cl = cr + 1;
        const uint32_t ntr = (ic + 1) * (ny + 1) + jc + 1, ntl = ic * (ny + 1) + jc + 1;
        const uint32_t nbl = ic * (ny + 1) + jc, nbr = (ic + 1) * (ny + 1) + jc;
        const uint32_t fR = (3 * ny + 1) * (ic + 1) + jc;
        const uint32_t fT = (3 * ny + 1) * ic + 2 * jc + ny + 2;
        const uint32_t fL = (3 * ny + 1) * ic + jc;
        const uint32_t fB = (3 * ny + 1) * ic + 2 * jc + ny;
        const uint32_t fD = (3 * ny + 1) * ic + 2 * jc + ny + 1;
        uint32_t* a = &m.nodes_of_cell[3 * (size_t)cr];
        a[0] = nbr; a[1] = ntr; a[2] = nbl;
        uint32_t* b = &m.nodes_of_cell[3 * (size_t)cl];
        b[0] = ntl; b[1] = nbl; b[2] = ntr;
        uint32_t* fa = &m.faces_of_cell[3 * (size_t)cr];
        fa[0] = fB; fa[1] = fR; fa[2] = fD;
        uint32_t* fb2 = &m.faces_of_cell[3 * (size_t)cl];
        fb2[0] = fT; fb2[1] = fL; fb2[2] = fD;
        auto set = [&](uint32_t f, int32_t c0, int32_t c1, uint32_t na, uint32_t nb) {
            m.cells_of_face[2 * (size_t)f] = c0; m.cells_of_face[2 * (size_t)f + 1] = c1;
            m.nodes_of_face[2 * (size_t)f] = na; m.nodes_of_face[2 * (size_t)f + 1] = nb;
        };
        set(fR, cr, ic == nx - 1 ? -1 : (int32_t)(cl + 2 * ny), nbr, ntr);
        set(fT, cl, jc == ny - 1 ? -1 : (int32_t)(cl + 1), ntr, ntl);
        set(fL, cl, ic == 0 ? -1 : (int32_t)(cr - 2 * ny), ntl, nbl);
        set(fB, cr, jc == 0 ? -1 : (int32_t)(cr - 1), nbl, nbr);
        set(fD, cr, cl, nbl, ntr);
        if (ic == nx - 1) fr.push_back(fR); else inner[fR] = 1;
        if (jc == ny - 1) ft.push_back(fT); else inner[fT] = 1;
        if (ic == 0) fl.push_back(fL); else inner[fL] = 1;
        if (jc == 0) fb.push_back(fB); else inner[fB] = 1;
        inner[fD] = 1;
    }
    zones_from(m, inner, std::move(fr), std::move(ft), std::move(fl), std::move(fb));
    m.compute_geometry();
    return m;
}

// mesh.cpp:828-848: 8-degree ramp for x > 0.5 applied to the quad mesh.
Mesh make_wedge(uint32_t nx, uint32_t ny, double Lx, double Ly) {
    Mesh m = make_cartesian(nx, ny, Lx, Ly);
    m.type = MeshType::Wedge;
    const double th = 8 * M_PI / 180.0, xw = 0.5;
    for (uint32_t n = 0; n < m.n_nodes; ++n) {
        const double x = m.node_coords[2 * (size_t)n], y = m.node_coords[2 * (size_t)n + 1];
        if (x > xw) {
            const double yb = (x - xw) * std::tan(th);
            m.node_coords[2 * (size_t)n + 1] = (y / Ly) * (Ly - yb) + yb;
        }
    }
    m.compute_geometry();
    return m;
}

Mesh make_from_arrays(uint32_t n_nodes, const double* node_coords, uint32_t n_cells,
                      const uint32_t* off_noc, const uint32_t* noc, const uint32_t* foc,
                      uint32_t n_faces, const uint32_t* nof, const int32_t* cof, uint32_t n_zones,
                      const char* const* zone_names, const uint32_t* zone_off,
                      const uint32_t* zone_faces) {
    Mesh m;
    m.type = MeshType::File;
    m.n_nodes = n_nodes;
    m.n_cells = n_cells;
    m.n_faces = n_faces;
    m.n_faces_real = n_faces;
    m.node_coords.assign(node_coords, node_coords + 2 * (size_t)n_nodes);
    m.offsets_nodes_of_cell.assign(off_noc, off_noc + n_cells + 1);
    m.offsets_faces_of_cell = m.offsets_nodes_of_cell;
    const size_t nn = off_noc[n_cells];
    m.nodes_of_cell.assign(noc, noc + nn);
    m.faces_of_cell.assign(foc, foc + nn);
    m.nodes_of_face.assign(nof, nof + 2 * (size_t)n_faces);
    m.offsets_nodes_of_face.resize(n_faces + 1);
    for (uint32_t f = 0; f <= n_faces; ++f) m.offsets_nodes_of_face[f] = 2 * f;
    m.cells_of_face.assign(cof, cof + 2 * (size_t)n_faces);
    for (size_t k = 0; k < nn; ++k)
        if (noc[k] >= n_nodes || foc[k] >= n_faces) fail("mesh arrays: index out of range");
    for (size_t k = 0; k < 2 * (size_t)n_faces; ++k) {
        if (nof[k] >= n_nodes) fail("mesh arrays: face node out of range");
        if (cof[k] < -1 || cof[k] >= (int64_t)n_cells) fail("mesh arrays: face cell out of range");
    }
    for (uint32_t z = 0; z < n_zones; ++z) {
        FaceZone Z;
        Z.name = zone_names[z];
        Z.interior = Z.name == "interior";
        for (uint32_t k = zone_off[z]; k < zone_off[z + 1]; ++k) {
            if (zone_faces[k] >= n_faces) fail("mesh arrays: zone face out of range");
            Z.faces.push_back(zone_faces[k]);
        }
        m.zones.push_back(std::move(Z));
    }
    if (m.zone_index("interior") != 0) fail("mesh arrays: zone 0 must be named 'interior'");
    m.compute_geometry();
    return m;
}

}  // namespace mlr
