// Shared host-side definitions for the MI355X residual engine.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

namespace mlr {

constexpr int kDim = 2;    // common_typedef.h:19
constexpr int kCons = 4;   // common_typedef.h:20
constexpr int kPrim = 5;   // common_typedef.h:21

// Enumerations keep the reference's string keys (see the *_TYPES maps next to
// each reference class).
enum class MeshType { File, Cartesian, CartesianTri, Wedge };             // mesh.h:23-35
enum class Recon { FO, TENO };                                            // face_reconstruction.h:30-33
enum class BasisType { Monomial, Legendre };                              // basis.h:24-32
enum class Riemann { Rusanov = 0, HLL = 1, HLLC = 2, Roe = 3 };           // riemann_solver.h:21-31 (+Roe)
enum class Integrator { FE, RK4, SSPRK3 };                                // time_integrator.h:29-33
// boundary.h:22-37; WallNoSlip (adiabatic no-slip wall, viscous runs) has no reference counterpart
enum class BCType { Symmetry = 0, Extrapolation = 1, WallAdiabatic = 2, UPT = 3, POut = 4, WallNoSlip = 5 };
enum class TenoVariant { Reference = 0, Corrected = 1 };

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] inline void fail(const std::string& msg) { throw Error(msg); }

}  // namespace mlr
