// Solution output in the reference's VTU layout (data_writer.cpp:93-244):
// an UnstructuredGrid with Float64 cell arrays, 3-component Float64 points,
// UInt32 connectivity / offsets, UInt8 VTK_POLYGON types, all in one raw
// appended block whose arrays each start with a 4-byte length.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace mlr {

// [[write_data]] entry (data_writer.cpp:29-77)
struct WriteSpec {
    std::string prefix;
    uint32_t interval = 1;
    std::vector<std::string> variables;
};

// Names a writer may select (common_typedef.h:36-50, solver.cpp:336-350).
bool is_output_variable(const std::string& name);

// File name of a step: prefix + "_" + 6-digit zero-padded step + ".vtu".
std::string vtu_name(const std::string& prefix, uint64_t step);

// Writes one VTU. Arrays are in reference numbering: U[N][4] conserved,
// prim[N][5] (u, v, p, T, h), cfl[N].
void write_vtu(const std::string& path, const Mesh& m, const std::vector<std::string>& variables, const double* U,
               const double* prim, const double* cfl);

}  // namespace mlr
