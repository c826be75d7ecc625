// Launch interface of the HIP kernels (implemented in kernels.hip). Host code
// passes plain device pointers; no HIP types appear here except the stream.
#pragma once

#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace mlr::dev {

struct Physics {
    double gamma, R, cp, cv, p_min, p_max;
};

struct BCParam {
    int32_t type;  // BCType
    int32_t pad;
    double u0, u1, p, T, rho, e, h;
};

struct Geom {
    uint32_t N, S, Q;
    const double* volume;
    const int32_t* nb;
    const uint32_t* info;
    const double* nx;
    const double* ny;
    const double* area;
    const double* dxn;
    const double* qw;     // face quadrature weights (device copy, Q entries)
    const BCParam* bc;    // boundary conditions in input order (device copy)
};

struct Teno {
    int32_t K, M, variant;
    double basis0;
    // stencil members as positions in the 64-cell block's tile list
    const uint8_t* loc8;       // [G][M-1][N] when every tile has <= 256 cells
    const uint16_t* loc16;     // [G][M-1][N] otherwise (<= 1024 cells)
    const uint32_t* blk_off;   // [n_blocks + 1]
    const int32_t* blk_ids;    // tile lists, ascending cell ids
    const double* ta;
    const double* A;
    const double* fbasis;
    double OI[81];      // (K-1) x (K-1) block of the oscillation indicator, row-major
};

// Stage programs of the fused residual + SSP-RK update.
enum StageOp : int32_t {
    kFE = 0,        // Un <- Un + dt k            (out: Uout, then copied back)
    kSSP1 = 1,      // Uout <- Un + dt k ; W <- Un + (dt c0) k
    kSSP2 = 2,      // Uout <- (3/4 Un + 1/4 Uin) + (dt/4) k ; W <- W + (dt c1) k
    kSSP3 = 3,      // Un <- W + (dt c2) k
    kRK4_1 = 4,     // Uout <- Un + (dt/2) k ; W <- Un + (dt c0) k
    kRK4_2 = 5,     // Uout <- Un + (dt/2) k ; W <- W + (dt c1) k
    kRK4_3 = 6,     // Uout <- Un + dt k     ; W <- W + (dt c2) k
    kRK4_4 = 7,     // Un <- W + (dt c3) k
    kRhsOnly = 8,   // rhs <- k
};

struct StageArgs {
    int32_t op;
    int32_t riemann;
    int32_t fo;           // 1: face states are the cell averages
    int32_t vacuum_guard;
    const double* Uin;    // stage input state (gathered), [N][4]
    const double* Un;     // step start state
    double* Uout;
    double* W;            // SSP-RK accumulator
    double* prims;        // [N][4] (u, v, p, T) written on the final stage
    double* h;            // [N] enthalpy written on the final stage
    double* rhs;          // kRhsOnly output
    const double* fs;     // face states [S*Q][N][4] (TENO)
    const double* dt;     // device scalar
};

// Launchers (asynchronous on `stream`). Each returns a hipError_t.
hipError_t teno_faces(const Geom& g, const Teno& t, const double* U, double* fs, uint8_t* decisions,
                      const Physics& ph, hipStream_t stream);
hipError_t flux_update(const Geom& g, const StageArgs& a, const Physics& ph, hipStream_t stream);
hipError_t spectral_radius(const Geom& g, const double* Un, const double* prims, const Physics& ph,
                           double* lam_cell, unsigned long long* lam_max, hipStream_t stream);
hipError_t finalize_dt(const unsigned long long* lam_max_in, double cfl, double fixed_dt, double* dt,
                       double* dt_hist, int64_t hist_cap, int64_t* step, unsigned long long* lam_max_reset,
                       hipStream_t stream);
hipError_t copy_state(const double* src, double* dst, uint32_t N, hipStream_t stream);
hipError_t probe_math(const double* x, const double* y, double* out, int n, int which, hipStream_t stream);
// Streaming bandwidth probes: kind 0 copies `bytes` from src to dst, kind 1
// reads them (dst receives stream_probe_threads(bytes) doubles).
uint32_t stream_probe_threads(size_t bytes);
hipError_t stream_probe(int kind, const void* src, void* dst, size_t bytes, hipStream_t stream);

}  // namespace mlr::dev
