// Device data layout: cells renumbered along a Hilbert curve, structure-of-
// arrays face-slot geometry, and the TENO stencil data in stencil-major,
// cell-minor order so that consecutive lanes read consecutive addresses.
//
// Numbering: "old" = reference numbering (mesh.cpp), "new" = SFC order.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "mesh.hpp"
#include "numerics.hpp"

namespace mlr {

struct BoundarySpec {
    std::string zone;
    BCType type = BCType::Symmetry;
    double u[2] = {0.0, 0.0};
    double p = 0.0, T = 0.0;
};

struct TenoSpec {
    BasisType basis = BasisType::Monomial;  // face_reconstruction.cpp:110 default
    int order = 1;
    double factor = 2.0;
    int quad_order_cell = -1;  // <0: reference default order + 1
    int quad_order_face = -1;  // <0: reference default (order + 1) / 2
    TenoVariant variant = TenoVariant::Reference;
};

// Slot info word (one per cell face slot):
//   bits 0-1  neighbour's slot of the shared face
//   bit  2    side: 1 when this cell is cells_of_face[f][1]
//   bit  3    boundary face
//   bits 4-7  boundary condition index (15: zone without a condition)
//   bits 8-15 (slot 0 only) accumulation order: rank r -> slot (2 bits each)
//   bits 16-18 (slot 0 only) number of slots that contribute to the residual
//   bits 20-23 (slot 0 only) bitmask of non-empty TENO stencils
//   bits 24-29 (slot 0 only, TENO) reference-triangle edge of each slot l
//              (bits 24 + 2 l: 0 nodes 0-1, 1 nodes 1-2, 2 nodes 0-2)
//   bits 30, 31, 19 (slot 0 only, TENO) set when slot 0, 1, 2 runs from the
//              edge's higher node to its lower one (dev::face_flip_shift)
constexpr uint32_t kSlotBoundary = 1u << 3;
constexpr uint32_t kNoBC = 15u;

struct Layout {
    uint32_t N = 0;        // cells
    uint32_t S = 0;        // face slots per cell (3 triangles, 4 quads)
    uint32_t Q = 1;        // face quadrature points
    std::vector<double> qface_w;  // face quadrature weights
    std::vector<double> qface_pts;  // face quadrature points on [-1, 1]
    std::vector<uint32_t> perm;   // new -> old
    std::vector<uint32_t> iperm;  // old -> new
    std::vector<uint32_t> face_slot;  // old face id -> (new cell << 2 | slot) of side 0, or ~0u

    std::vector<double> volume;           // [N]
    std::vector<double> centroid;         // [N][2] (new numbering, for init/output)
    std::vector<int32_t> nb;              // [S][N] neighbour (new id), self for boundary slots
    std::vector<uint32_t> info;           // [S][N]
    std::vector<double> nx, ny, area, dxn;  // [S][N]
    std::vector<double> dcx, dcy;           // [S][N] centroid to neighbour centroid (face midpoint on boundaries)
    std::vector<double> mfx, mfy;           // [S][N] centroid to face midpoint

    // TENO (empty for FO): G = S + 1 stencils per cell, K = n_dof, M = cells per stencil.
    bool teno = false;
    int K = 0, M = 0, G = 0;
    TenoSpec tspec;
    std::vector<uint8_t> poly;            // [K][2]
    std::vector<double> ipt;              // [K]
    std::vector<double> OI;               // [K][K]
    std::vector<int32_t> st_ids;          // neighbours in new numbering (mem_idx)
    std::vector<double> st_ta;            // transformed member areas (mem_idx)
    std::vector<double> st_A;             // A+ without its zero row/column (a_idx)
    std::vector<double> fbasis;           // psi_k(x_q) + ipt_k/ta_0 (or - ipt_k) (fb_idx)
    double basis0 = 1.0;                  // sign carrier of the k = 0 term

    // LDS tiles of the TENO kernel: block b = cells [64b, 64b+64) (new
    // numbering). blk_ids[blk_off[b] .. blk_off[b+1]) is the sorted list of
    // every cell the block's stencils read (its own cells included); st_loc
    // gives each stencil member's position in its block's list.
    static constexpr uint32_t kTileCells = 64;
    static constexpr uint32_t kTileMax8 = 256;    // list length that 8-bit positions address
    static constexpr uint32_t kTileMax16 = 1024;  // LDS cap of the 16-bit form
    std::vector<uint32_t> blk_off;        // [n_blocks + 1]
    std::vector<int32_t> blk_ids;         // [blk_off.back()]
    std::vector<uint8_t> st_loc8;         // (mem_idx), when every list has <= 256 entries
    std::vector<uint16_t> st_loc16;       // (mem_idx), otherwise
    uint32_t tile_max = 0;                // longest list
    // true: st_A / st_ta are left empty and computed on the GPU at upload
    bool device_aplus = false;
    std::vector<double> qcell_pts, qcell_w;  // cell quadrature (for the device precompute)

    uint64_t device_bytes() const;

    // Blocked (AoSoA) indices of the per-stencil tables: cells in blocks of
    // 64 (one tile), member-major inside a block, so the data one thread
    // block reads is one contiguous range, and for each member the block's
    // cells x stencils cell-major (G = 4), so a wave's 64 lanes (16 cells x 4
    // stencils) read 64 consecutive entries (dev::teno_mem_off, teno_a_off).
    // i counts the members after the target (0 .. M-2), fq = slot * Q + point.
    //   ids / ta / positions: [N/64][M-1][64][G]
    //   A+:                   [N/64][M-1][K-1][64][G]
    //   face basis:           [N/64][S*Q][K-1][64]
    size_t n_pad() const { return ((size_t)N + 63) / 64 * 64; }
    size_t mem_idx(int s, int i, uint32_t c) const { return dev::teno_mem_off(c >> 6, M - 1, i, c & 63u, s); }
    size_t a_idx(int s, int i, int k, uint32_t c) const {
        return dev::teno_a_off(c >> 6, M - 1, K - 1, i, k, c & 63u, s);
    }
    size_t fb_idx(uint32_t fq, int k, uint32_t c) const {
        return (((size_t)(c >> 6) * (S * Q) + fq) * (K - 1) + k) * 64 + (c & 63u);
    }
};
static_assert(Layout::kTileCells == 64, "the blocked tables assume 64-cell blocks");

// Hilbert-curve order of cell centroids (ties broken by reference id).
std::vector<uint32_t> sfc_order(const Mesh& m);

// Builds the layout. bcs lists the boundary conditions in input order.
// allow_device: leave the pseudo-inverses to the GPU precompute when it
// supports the basis and stencil size (the host still searches stencils).
Layout build_layout(const Mesh& m, Recon recon, const TenoSpec& ts,
                    const std::vector<BoundarySpec>& bcs, int n_threads = 0, bool allow_device = false);

// Reference-numbered stencil lists recomputed for one cell (used by the
// index-parity tests through the C-ABI): out = G stencils of M ids each (or
// empty), in reference numbering.
void stencils_of_cell(const Mesh& m, int M, uint32_t c, std::vector<std::vector<uint32_t>>& out);

// Full K x M pseudo-inverses of a cell's stencils in reference order
// (K*M doubles per non-empty stencil), for parity tests.
void teno_matrices_of_cell(const Mesh& m, const TenoSpec& ts, uint32_t c,
                           std::vector<std::vector<double>>& aplus,
                           std::vector<std::vector<double>>& ta, std::vector<double>& ipt_out);

}  // namespace mlr
