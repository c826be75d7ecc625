#include "io.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "common.hpp"

namespace mlr {

namespace {

const char* const kNames[] = {"RHO", "RHOU_X", "RHOU_Y", "RHOE", "U_X", "U_Y", "P", "T", "H", "CFL"};

int var_index(const std::string& n) {
    for (int i = 0; i < 10; ++i)
        if (n == kNames[i]) return i;
    return -1;
}

bool little_endian() {
    const uint16_t one = 1;
    uint8_t b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

// 4-byte length header of an appended array (the low 32 bits of its size)
void put_len(std::ostream& o, uint64_t bytes) {
    const uint32_t n = (uint32_t)bytes;
    o.write(reinterpret_cast<const char*>(&n), 4);
}

}  // namespace

bool is_output_variable(const std::string& name) { return var_index(name) >= 0; }

std::string vtu_name(const std::string& prefix, uint64_t step) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "_%06llu.vtu", (unsigned long long)step);
    return prefix + buf;
}

void write_vtu(const std::string& path, const Mesh& m, const std::vector<std::string>& variables, const double* U,
               const double* prim, const double* cfl) {
    const uint32_t N = m.n_cells, Nn = m.n_nodes;
    std::vector<int> vi;
    for (const auto& v : variables) {
        const int i = var_index(v);
        if (i < 0) fail("DataWriter: Unknown variable: " + v + ".");
        vi.push_back(i);
    }
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) fail("DataWriter::write_vtu: Could not open file: " + path + ".");

    uint64_t len_conn = 0;
    for (uint32_t c = 0; c < N; ++c) len_conn += m.n_nodes_of_cell(c);

    std::ostringstream h;
    uint64_t off = 0;
    h << "<?xml version=\"1.0\"?>\n";
    h << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\""
      << (little_endian() ? "LittleEndian" : "BigEndian") << "\">\n";
    h << "  <UnstructuredGrid>\n";
    h << "    <Piece NumberOfPoints=\"" << Nn << "\" NumberOfCells=\"" << N << "\">\n";
    h << "      <PointData>\n      </PointData>\n";
    h << "      <CellData>\n";
    for (const auto& v : variables) {
        h << "        <DataArray type=\"Float64\" Name=\"" << v << "\" format=\"appended\" offset=\"" << off
          << "\">\n        </DataArray>\n";
        off += 4 + (uint64_t)N * 8;
    }
    h << "      </CellData>\n";
    h << "      <Points>\n";
    h << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << off
      << "\">\n        </DataArray>\n";
    off += 4 + (uint64_t)Nn * 3 * 8;
    h << "      </Points>\n";
    h << "      <Cells>\n";
    h << "        <DataArray type=\"UInt32\" Name=\"connectivity\" format=\"appended\" offset=\"" << off
      << "\">\n        </DataArray>\n";
    off += 4 + len_conn * 4;
    h << "        <DataArray type=\"UInt32\" Name=\"offsets\" format=\"appended\" offset=\"" << off
      << "\">\n        </DataArray>\n";
    off += 4 + (uint64_t)N * 4;
    h << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" << off
      << "\">\n        </DataArray>\n";
    h << "      </Cells>\n";
    h << "    </Piece>\n";
    h << "  </UnstructuredGrid>\n";
    h << "<AppendedData encoding=\"raw\">\n_";
    const std::string head = h.str();
    out.write(head.data(), (std::streamsize)head.size());

    std::vector<double> col(N);
    for (int i : vi) {
        for (uint32_t c = 0; c < N; ++c)
            col[c] = i < 4 ? U[4 * (size_t)c + i] : (i < 9 ? prim[5 * (size_t)c + (i - 4)] : cfl[c]);
        put_len(out, (uint64_t)N * 8);
        out.write(reinterpret_cast<const char*>(col.data()), (std::streamsize)N * 8);
    }
    std::vector<double> pts(3 * (size_t)Nn);
    for (uint32_t n = 0; n < Nn; ++n) {
        pts[3 * (size_t)n] = m.nc(n, 0);
        pts[3 * (size_t)n + 1] = m.nc(n, 1);
        pts[3 * (size_t)n + 2] = 0.0;
    }
    put_len(out, (uint64_t)Nn * 3 * 8);
    out.write(reinterpret_cast<const char*>(pts.data()), (std::streamsize)pts.size() * 8);
    std::vector<uint32_t> conn;
    conn.reserve(len_conn);
    std::vector<uint32_t> offs(N);
    uint32_t acc = 0;
    for (uint32_t c = 0; c < N; ++c) {
        for (uint32_t l = 0; l < m.n_nodes_of_cell(c); ++l) conn.push_back(m.node_of_cell(c, (int)l));
        acc += m.n_nodes_of_cell(c);
        offs[c] = acc;
    }
    put_len(out, len_conn * 4);
    out.write(reinterpret_cast<const char*>(conn.data()), (std::streamsize)conn.size() * 4);
    put_len(out, (uint64_t)N * 4);
    out.write(reinterpret_cast<const char*>(offs.data()), (std::streamsize)N * 4);
    put_len(out, (uint64_t)N);
    const std::vector<uint8_t> types(N, 7);  // VTK_POLYGON
    out.write(reinterpret_cast<const char*>(types.data()), (std::streamsize)N);
    out << "\n</AppendedData>\n</VTKFile>\n";
    if (!out.good()) fail("DataWriter::write_vtu: Could not write to file: " + path + ".");
}

}  // namespace mlr
