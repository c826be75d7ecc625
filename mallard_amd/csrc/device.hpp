// Launch interface of the HIP kernels (implemented in kernels.hip). Host code
// passes plain device pointers; no HIP types appear here except the stream.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace mlr::dev {

// Blocked TENO member tables (stencil ids, tile positions, member areas; A+
// with one more level for its dof k): for 64-cell block b and member i
// after the target, the block's 64 cells x 4 stencils, cell-major. A wave's
// 64 lanes (16 cells x 4 stencils, lane = 4 cell + stencil) then read 64
// consecutive entries for each (i, k).
constexpr size_t teno_mem_off(size_t b, int Mp, int i, uint32_t cl, int s) {
    return ((b * (size_t)Mp + (size_t)i) * 64 + cl) * 4 + (size_t)s;
}
constexpr size_t teno_a_off(size_t b, int Mp, int KP, int i, int k, uint32_t cl, int s) {
    return (((b * (size_t)Mp + (size_t)i) * (size_t)KP + (size_t)k) * 64 + cl) * 4 + (size_t)s;
}
// the central stencil's linear operator of gated runs: for block b and
// member i, the block's 64 cells x FQ face points, cell-major
constexpr size_t teno_lin_off(size_t b, int Mp, int i, uint32_t cl, int FQ, int fq) {
    return ((b * (size_t)Mp + (size_t)i) * 64 + cl) * (size_t)FQ + (size_t)fq;
}
// Bit of the slot-0 info word that is set when face slot l (0..2) of a
// triangle runs from its edge's higher cell node to its lower one (the
// face's node 0 is the cell node hi of the edge (lo, hi) in bits 24 + 2 l)
constexpr int face_flip_shift(int l) { return l == 0 ? 30 : (l == 1 ? 31 : 19); }

struct Physics {
    double gamma, R, cp, cv, p_min, p_max;
    // Navier-Stokes: viscous != 0 adds the viscous fluxes; Sutherland law
    // mu = mu_ref (T/T_mu)^1.5 (T_mu + S)/(T + S) (constant mu_ref when S < 0)
    int32_t viscous;
    double mu_ref, T_mu, S_mu, Pr;
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
    const double* dcx;    // [S][N] centroid to neighbour centroid, or to the face
    const double* dcy;    //        midpoint on boundaries (viscous runs only)
    const double* mfx;    // [S][N] centroid to face midpoint (viscous runs only)
    const double* mfy;
};

namespace ind {
constexpr int kNParams = 7361;  // 48*64 + 64 + 64*64 + 64 + 64 + 1
// Device copy of the indicator network as the kernels read it: W1p / W2p are
// the bf16 MFMA operand layouts, the rest FP32.
struct Net {
    const void* W1p;
    const void* W2p;
    const float* b1;
    const float* b2;
    const float* w3;
    const float* b3;
    // a cell is judged smooth when sigmoid(z) >= tau, that is z >= log(tau / (1 - tau))
    // (+inf for tau = 1: never)
    float logit_threshold;
};
}  // namespace ind

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
    // learned indicator gate: when gate != 0 a cell the network judges smooth
    // takes the central stencil alone; gated counts smooth cells, one
    // running total per wave slot (block * 4 + wave), each written only by
    // its own wave, so the count needs no atomics
    int32_t gate;
    ind::Net net;
    uint32_t* gated;
    uint8_t* gate_dec;  // per cell: 1 when the last network evaluation judged it smooth
    // gated runs: the central stencil's linear reconstruction as one operator
    // (teno_lin_off): face point fq of cell c is
    // u_c + sum_i lin(fq, i) (u_i - u_c) over the stencil's members i
    const double* lin;
    double gamma;
    // non-null: gated passes run split (teno_kernel modes 3/4, then 5 for
    // the ungated cells); one byte per wave slot (block * 4 + wave), 1 when
    // the wave holds a valid cell the network did not gate
    uint8_t* blk_mixed;
    // compact tables (corrected weighting, one face point per face): A holds
    // A+ times the member areas (ta unused) and the face-point basis values
    // come from fbc[e][K-1], one row per reference-triangle edge e (the
    // midpoint of nodes 0-1, 1-2 or 2-0), e of slot l in bits 24 + 2 l of the
    // slot-0 info word
    int32_t compact;
    const double* fbc;
    // derived face basis (p = 2, Legendre basis, the default stencil size):
    // the plain pass recomputes the face-point basis values from the cell's
    // node coordinates with the precompute's own operations (bit-identical
    // to fbasis, which it then does not read). cvert: [N][6] the cell's
    // nodes (x0, y0, x1, y1, x2, y2) in its node order; ipt, poly: the
    // basis' cell means and (px, py) per dof; qf: the face quadrature point
    // on [-1, 1] (one per face)
    int32_t derived;
    const double* cvert;
    double ipt[10];
    uint8_t poly[20];
    double qf;
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
    // 1: the corrected TENO variant's flux rules: the vacuum guard of the
    // two-rarefaction estimate, and extrapolation boundaries whose ghost state
    // is the cell mean (a zero-gradient ghost cell) instead of the face state
    int32_t corrected;
    int32_t pp;           // 1: positivity-preserving flux limiter (stage updates)
    const double* Uin;    // stage input state (gathered), [N][4]
    const double* Un;     // step start state
    double* Uout;
    double* W;            // SSP-RK accumulator
    double* prims;        // [N][4] (u, v, p, T) written on the final stage
    double* h;            // [N] enthalpy written on the final stage
    double* rhs;          // kRhsOnly output
    const double* fs;     // face states [S*Q][N][4] (TENO)
    const double* dt;     // device scalar
    const double* vp;     // viscous runs: (u, v, T, mu) [N][4] and the
    const double* vg;     // cell gradients (u_x, u_y, v_x, v_y, T_x, T_y) [N][6] of Uin
};

// One case of a batched ensemble: its geometry, stencils and buffers, and the
// first block of the case in the batched TENO grid (64-cell blocks) and in the
// flux / spectral grids (256-cell blocks).
struct CaseBufs {
    Geom g;
    Teno t;
    double *Un, *Ua, *Ub, *W, *prims, *h, *fs, *rhs, *lam_cell, *dt, *dt_hist;
    unsigned long long* lam_blk;  // per-block spectral maxima, room for (g.N + 63) / 64 of them
    int64_t* step;
    uint8_t* dec;
    uint32_t blk_teno0, blk_flux0;
};
enum BufRole : int32_t { kBufUn = 0, kBufUa = 1, kBufUb = 2, kBufNone = 3 };
struct BatchStage {
    int32_t op, fo, corrected;
    int32_t in, out;  // BufRole of the stage input and output
    int32_t pp = 0;   // 1: positivity-preserving flux limiter
};

// Launch shapes an autotuner picks per mesh. Every choice gives identical
// results; only the speed differs.
struct Tune {
    int32_t teno_occ = 0;       // TENO waves per SIMD targeted with stencil factor 2: 0 (the
                                // compiler's choice); 5 or 6 at p = 2, 4 or 5 for plain p = 3
    uint32_t flux_block = 256;  // cells per flux / stage-update block: 64, 128 or 256
    uint32_t spec_block = 256;  // cells per spectral-radius block: 64, 128 or 256
    int32_t teno_occ_reuse = 0; // the same target for gated passes that reuse recorded decisions
    int32_t teno_face_basis = 0; // plain TENO pass: 0 reads the face-basis table, 1 derives
                                // it from the cell's node coordinates (Teno::derived)
};

// Launchers (asynchronous on `stream`). Each returns a hipError_t.
// gate_reuse: with the gate on, take each cell's recorded decision instead of
// evaluating the network again (later stages of a step)
hipError_t teno_faces(const Geom& g, const Teno& t, const double* U, double* fs, uint8_t* decisions,
                      const Physics& ph, int occ, int gate_reuse, hipStream_t stream);
// bs: cells (threads) per block
hipError_t flux_update(const Geom& g, const StageArgs& a, const Physics& ph, uint32_t bs, hipStream_t stream);
// kRhsOnly residual over the faces `parts` selects: 1 interior, 2 boundary,
// 3 both (the full residual, flux_update)
hipError_t rhs_parts(const Geom& g, const StageArgs& a, const Physics& ph, int parts, uint32_t bs,
                     hipStream_t stream);
// spectral radii: lam_cell per cell and, per bs-cell block, the bit pattern
// of the block's largest radius in lam_blk (NaN radii are ignored)
hipError_t spectral_radius(const Geom& g, const double* Un, const double* prims, const Physics& ph,
                           double* lam_cell, unsigned long long* lam_blk, uint32_t bs, hipStream_t stream);
// dt = cfl / (max over the n_blk block maxima), or fixed_dt when > 0
hipError_t finalize_dt(const unsigned long long* lam_blk, uint32_t n_blk, double cfl, double fixed_dt, double* dt,
                       double* dt_hist, int64_t hist_cap, int64_t* step, hipStream_t stream);
hipError_t copy_state(const double* src, double* dst, uint32_t N, hipStream_t stream);
// viscous runs: per cell (u, v, T, mu) and least-squares gradients of u, v, T
// of the state U (bs cells per block)
hipError_t viscous_grad(const Geom& g, const double* U, const Physics& ph, double* vp, double* vg, uint32_t bs,
                        hipStream_t stream);
// Ensemble launchers: `cases` is a device array of n_cases descriptors whose
// blk_* offsets are prefix sums (flux and spectral blocks of bs cells, TENO
// blocks of 64); every case has the same K, M, face points, tile-position
// width (small = 8-bit) and Riemann solver.
// gated: some case runs the learned gate (the block stages the network in LDS);
// split: every case's gate runs split (Teno::blk_mixed set), as teno_faces
// derived: plain pass with every case's face basis derived (Teno::derived)
hipError_t teno_batch(const CaseBufs* cases, int n_cases, uint32_t total_blocks, int in_role, int K, int M, int FQ,
                      bool small, bool gated, int occ, int gate_reuse, bool split, bool derived, const Physics& ph,
                      hipStream_t stream);
hipError_t flux_batch(const CaseBufs* cases, int n_cases, uint32_t total_blocks, const BatchStage& st, int riemann,
                      const Physics& ph, uint32_t bs, hipStream_t stream);
hipError_t spectral_batch(const CaseBufs* cases, int n_cases, uint32_t total_blocks, const Physics& ph, uint32_t bs,
                          hipStream_t stream);
hipError_t finalize_batch(const CaseBufs* cases, int n_cases, double cfl, double fixed_dt, int64_t hist_cap,
                          uint32_t bs, hipStream_t stream);
// A whole ensemble step in one persistent cooperative launch (the spectral
// radii and dt of every case, then per stage the TENO and flux + update
// phases with grid barriers between them): plain TENO at p = 2 with the
// default stencil size and 8-bit tile positions (teno) or first order, flux
// and spectral blocks of 256 cells in the case table. Stage st runs op[st]
// from BufRole in[st] to out[st]. grid_out receives the blocks launched.
// Barrier state of the persistent step (zeroed once; each barrier leaves
// count at 0). err: a block gave up waiting.
struct StepBarrier {
    uint32_t count, gen, err, pad;
};
struct PersistentStep {
    int32_t n_stages;
    int32_t op[4], in[4], out[4];
    int32_t fo, corrected, use_cfl;
    double cfl, fixed_dt;
    int64_t hist_cap;
    // NULL, or 3 + 2 n_stages device wall-clock stamps taken by block 0 after
    // each barrier (start, radii, dt, then per stage TENO and flux): the
    // flux phase of the last stage then ends with a barrier too
    long long* clocks;
    StepBarrier* bar;
    long long spin_ticks;  // longest wait at a barrier, device wall-clock ticks
};
hipError_t ensemble_step(const CaseBufs* cases, int n_cases, uint32_t teno_total, uint32_t flux_total,
                         const PersistentStep& ps, int riemann, bool teno, const Physics& ph, hipStream_t stream,
                         int* grid_out);
hipError_t probe_math(const double* x, const double* y, double* out, int n, int which, hipStream_t stream);

// Point-wise operators (operators.hip).
// RiemannSolver::calc_flux of n face points: nrm[n][2] unit normals,
// L, R[n][6] = (rho, u, v, p, gamma, h) left / right, F[n][4] fluxes; type is
// the Riemann solver code, guard the vacuum guard of the two-rarefaction
// estimate (0: the reference's formula)
hipError_t riemann_points(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R,
                          double* F, hipStream_t stream);
// vector updates of n doubles: kind 0 y = x (deep_copy), 1 y = y + a x
// (axpy), 2 y = a x + b y (axpby)
hipError_t vec_update(int kind, int64_t n, double a, const double* x, double b, double* y, hipStream_t stream);
// The Riemann operator on HOST arrays (same layouts): stages the points
// through device memory allocated for the call, on the null stream; throws
// mlr::Error on failure
void riemann_flux_host(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R,
                       double* F);
// The same without allocating: the points are staged through a fixed device
// array of riemann_flux_noalloc_capacity() points; larger batches are refused.
// Calls that overlap on one device share that array and race.
void riemann_flux_noalloc(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R,
                          double* F);
int64_t riemann_flux_noalloc_capacity();

// TENO pseudo-inverses on the GPU (precompute.hip), bit-identical to the host
// precompute for the Legendre basis.
struct PreInput {
    uint32_t N;
    const double* xy;      // node coordinates [n_nodes][2]
    const uint32_t* tri;   // nodes of each cell, new numbering, reference node order [N][3]
    const int32_t* ids;    // stencil members (new numbering), blocked like the A+ table
    const uint32_t* info;  // slot-0 info words (stencil mask in bits 20-23)
    int32_t nq;            // cell quadrature points (<= 8)
    double qp[16], qw[8];
    uint8_t poly[20];      // [K][2] Legendre degrees
    double ipt[10];
};
bool aplus_supported(int K, int M);
// lin(fq, i) = (sum_k fbasis(fq, k) A+(central, i, k)) ta(central, i), 0 where
// a cell has no central stencil (the blocked layouts of the Teno tables)
hipError_t teno_linear_operator(const Teno& t, const uint32_t* info, uint32_t N, int FQ, double* lin,
                                hipStream_t stream);
// A: [N/64][4][M-1][K-1][64], ta: [N/64][4][M-1][64]; err |= 1 singular
// frame, |= 2 a stencil needing the full-matrix branch
hipError_t teno_aplus(const PreInput& in, int K, int M, double* A, double* ta, int* err, hipStream_t stream);
// A[row][k] *= ta[row] over the blocked tables (rows = n_blocks * G * (M-1)):
// the member areas folded into the pseudo-inverses (compact tables)
hipError_t teno_fold_areas(double* A, const double* ta, uint64_t rows, int KP, hipStream_t stream);

// Learned indicator (indicator.hip). The network lives in one device buffer of
// net_bytes() that net_pack fills from the FP32 master parameters.
size_t net_bytes();
ind::Net net_view(void* buf, double threshold, const float** W2f);
hipError_t net_pack(const float* theta, void* buf, hipStream_t stream);
// features (bf16 [N][48]) and TENO labels of every cell at rows base..base+N-1
hipError_t ind_features(const Geom& g, const Teno& t, const double* U, const uint8_t* dec, double gamma,
                        uint16_t* X, float* y, uint64_t base, hipStream_t stream);
hipError_t ind_infer(const ind::Net& net, const uint16_t* X, uint64_t n, float* z, hipStream_t stream);
int ind_train_blocks(uint32_t B);
// gradient of the weighted mean BCE over rows idx[0..B) (part: ind_train_blocks(B) x kNParams)
hipError_t ind_grad(const ind::Net& net, const float* W2f, const uint16_t* X, const float* y, const uint32_t* idx,
                    uint32_t B, float pos_w, float neg_w, float* part, float* loss_part, float* grad, float* loss,
                    hipStream_t stream);
hipError_t ind_adam(float* theta, float* m, float* v, const float* grad, float lr, int step, hipStream_t stream);
hipError_t ind_batch(const uint32_t* pool, uint64_t n_pool, uint64_t seed, uint64_t step, uint32_t B,
                     uint32_t* idx, hipStream_t stream);
// Streaming bandwidth probes: kind 0 copies `bytes` from src to dst, kind 1
// reads them (dst receives stream_probe_threads(bytes) doubles).
uint32_t stream_probe_threads(size_t bytes);
hipError_t stream_probe(int kind, const void* src, void* dst, size_t bytes, hipStream_t stream);

}  // namespace mlr::dev
