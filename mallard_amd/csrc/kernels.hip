// HIP kernels of the MI355X residual engine (gfx950, wave64, FP64).
//
// Numerics follow the reference operation by operation so that, compiled
// with -ffp-contract=off, every result matches the CPU oracle bit for bit:
//   - TENO reconstruction: face_reconstruction.cpp:866-1039
//   - Riemann fluxes:      riemann_solver.h:205-519
//   - flux functors / BCs: flux_functor.h:124-178, boundary_*.cpp
//   - residual assembly:   solver_rhs.cpp:44-55 (canonical one-thread order)
//   - SSP-RK / RK4 / FE:   time_integrator.cpp:57-163
//   - spectral radius, dt: solver.cpp:580-742
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ddpow.hpp"
#include "device.hpp"
#include "indicator.hpp"
#include "riemann.hpp"

namespace mlr::dev {

namespace {

constexpr uint32_t kBoundaryBit = 1u << 3;


struct D4 {
    double v[4];
};

__device__ inline D4 ld4(const double* __restrict__ p, size_t i) {
    const double2* q = reinterpret_cast<const double2*>(p + 4 * i);
    const double2 a = q[0], b = q[1];
    return D4{{a.x, a.y, b.x, b.y}};
}
__device__ inline void st4(double* __restrict__ p, size_t i, const D4& x) {
    double2* q = reinterpret_cast<double2*>(p + 4 * i);
    q[0] = make_double2(x.v[0], x.v[1]);
    q[1] = make_double2(x.v[2], x.v[3]);
}

// Correctly rounded x^6 (double-double product, exact FMA residuals); equals
// glibc's pow(x, 6.0) wherever that is correctly rounded.
__device__ inline double pow6_cr(double x) {
    const double h2 = x * x;
    const double l2 = __builtin_fma(x, x, -h2);
    const double p3 = h2 * x;
    const double e3 = __builtin_fma(h2, x, -p3) + l2 * x;
    const double h3 = p3 + e3;
    const double l3 = e3 - (h3 - p3);
    const double p6 = h3 * h3;
    const double e6 = __builtin_fma(h3, h3, -p6) + 2.0 * h3 * l3;
    return p6 + e6;
}

__device__ inline double pressure_raw(double gamma, const double* u) {
    return (gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
}

// ------------------------------------------------------------ quad exchange
// The 4 lanes of a quad serve one cell. DPP quad permutations move values
// between them in registers (no LDS, so occupancy is set by VGPRs alone).
template <int CTRL>
__device__ inline double qperm(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// quad_perm controls: lane j of a quad reads lane (j - r) & 3, r = 1, 2, 3;
// (kRot2 is also lane j ^ 2)
constexpr int kRot1 = 0x93, kRot2 = 0x4E, kRot3 = 0x39;

__device__ inline double sel4(const double* x, int i) {
    return i == 0 ? x[0] : (i == 1 ? x[1] : (i == 2 ? x[2] : x[3]));
}
// 4 x 4 transpose across a quad: out[j] = x[q] of quad lane j (q = own lane).
__device__ inline void quad_transpose(const double* x, double* out, int q) {
    double r[4];
    r[0] = sel4(x, q);
    r[1] = qperm<kRot1>(sel4(x, (q + 1) & 3));
    r[2] = qperm<kRot2>(sel4(x, (q + 2) & 3));
    r[3] = qperm<kRot3>(sel4(x, (q + 3) & 3));
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = sel4(r, (q - j) & 3);
}
// out[j] = f of quad lane j.
__device__ inline void quad_gather(double f, double* out, int q) {
    double r[4];
    r[0] = f;
    r[1] = qperm<kRot1>(f);
    r[2] = qperm<kRot2>(f);
    r[3] = qperm<kRot3>(f);
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = sel4(r, (q - j) & 3);
}

// ------------------------------------------------------------ derived tables
// The target cell's frame: node 0 and the inverse of the Jacobian of its
// map from the reference triangle, as the precompute forms them.
struct TriFrame {
    double w0x, w0y, j0, j1, j2, j3;
};
__device__ inline TriFrame tri_frame(double x0, double y0, double x1, double y1, double x2, double y2) {
    TriFrame F;
    F.w0x = x0;
    F.w0y = y0;
    const double J0 = x1 - x0, J1 = x2 - x0, J2 = y1 - y0, J3 = y2 - y0;
    const double det = J0 * J3 - J1 * J2;
    const double id = 1.0 / det;
    F.j0 = J3 * id;
    F.j1 = -J1 * id;
    F.j2 = -J2 * id;
    F.j3 = J0 * id;
    return F;
}
// a point in the frame: Ji (x - w0)
__device__ inline void to_frame(const TriFrame& F, double x, double y, double& a, double& b) {
    const double tx = x - F.w0x, ty = y - F.w0y;
    a = F.j0 * tx + F.j1 * ty;
    b = F.j2 * tx + F.j3 * ty;
}
// area in the frame of the triangle with nodes (x0, y0), (x1, y1), (x2, y2)
__device__ inline double frame_area(const TriFrame& F, double x0, double y0, double x1, double y1, double x2,
                                    double y2) {
    double a0, b0, a1, b1, a2, b2;
    to_frame(F, x0, y0, a0, b0);
    to_frame(F, x1, y1, a1, b1);
    to_frame(F, x2, y2, a2, b2);
    return 0.5 * fabs(a0 * (b1 - b2) + a1 * (b2 - b0) + a2 * (b0 - b1));
}
// Legendre P_p(x) for p <= 2, terms combined as the basis evaluator does:
// P_2 = 0.5 ((3 x) x - 1)
__device__ inline double legendre_le2(int p, double x) {
    const double p2 = 0.5 * ((3.0 * x) * x - 1.0);
    return p == 0 ? 1.0 : (p == 1 ? x : p2);
}

// ------------------------------------------------------------ TENO kernel
// 64 cells per 256-thread block, 4 lanes (a quad) per cell.
// Phase A: quad lane s owns stencil s: stencil-major loads of ids, areas and
// A+, gathers of the members' states, a = A+ b, smoothness, raw weight.
// Quad transposes then hand lane v the 4 stencils' coefficients and weights
// of conserved variable v. Phase B: cut-off, weights and the face-point sums
// in the reference's stencil-then-dof order; with the corrected weighting the
// quad then scales its face states toward the cell mean for positivity.
// KP = n_dof - 1, MP = members per stencil after the target (0: runtime),
// FQ = face points per cell; IdxT is the width of the in-tile positions and
// TILE the LDS capacity in cells. The block first stages the states of every
// cell its stencils read (the tile list) in LDS, so each neighbour state is
// fetched once per block instead of once per stencil member. Blocks are
// mapped so that each of the 8 XCDs walks one contiguous range of the
// space-filling-curve order.
// GM: 0 plain TENO; 1 the learned gate evaluates the network and records
// its decisions; 2 the gate takes the recorded decisions (no network code).
// Split passes (no full-TENO code in the gated kernels, so they keep few
// registers): 3 evaluates like 1 and 4 reuses like 2, but both reconstruct
// only the gated cells and mark the waves holding ungated ones; 5 then
// reconstructs only the ungated cells, in the marked blocks.
// CT: compact tables (corrected weighting): A carries the member areas and
// the face-point basis values come from the small per-edge table t.fbc, so
// neither ta nor the per-cell face basis is read.
// DV: derived face basis (p = 2, Legendre basis, one face point per face):
// the face-point basis values are recomputed from the cell's node
// coordinates with the precompute's operations in its order, so they equal
// the table bit for bit, and the per-cell face basis is not read.
template <int KP, int MP, int FQ, typename IdxT, int TILE, int GM, bool CT = false, bool DV = false>
__device__ __forceinline__ void teno_block(const Geom& g, const Teno& t, const double* __restrict__ U,
                                           double* __restrict__ fs, uint8_t* __restrict__ dec, const Physics& ph,
                                           uint32_t lb, double2* sh_u, unsigned char* sh_net) {
    constexpr bool kEval = GM == 1 || GM == 3;
    constexpr bool kReuse = GM == 2 || GM == 4 || GM == 5;
    constexpr bool kLinOnly = GM == 3 || GM == 4;
    constexpr bool kFullOnly = GM == 5;
    const uint32_t N = g.N;
    if (lb * 64u >= N) return;  // whole block past the end (uniform)
    if constexpr (kFullOnly) {
        // the 4 wave marks of the block as one word (uniform)
        if (reinterpret_cast<const uint32_t*>(t.blk_mixed)[lb] == 0u) return;
    }
    {
        const uint32_t t0 = t.blk_off[lb], nloc = t.blk_off[lb + 1] - t0;
        const double2* U2 = reinterpret_cast<const double2*>(U);
        for (uint32_t j = threadIdx.x; j < 2 * nloc; j += blockDim.x)
            sh_u[j] = U2[2 * (size_t)t.blk_ids[t0 + (j >> 1)] + (j & 1u)];
        if (kEval && t.gate) ind::stage_net(t.net, sh_net);  // the network, once per block
    }
    __syncthreads();
    const IdxT* __restrict__ loc = sizeof(IdxT) == 1 ? reinterpret_cast<const IdxT*>(t.loc8)
                                                     : reinterpret_cast<const IdxT*>(t.loc16);
    const int q = threadIdx.x & 3;
    const uint32_t c = lb * 64u + (threadIdx.x >> 2);
    const bool valid = c < N;
    const uint32_t cc = valid ? c : 0;
    const int Mp = MP > 0 ? MP : t.M - 1;

    // ---- learned gate: the MLP judges each cell from its central stencil
    // (features from the LDS tile, one MFMA column per cell); a smooth cell
    // skips its directional stencils and takes the central one.
    bool smooth = false;
    if (kReuse && t.gate) {
        // the decisions the network recorded earlier in the step
        smooth = valid && t.gate_dec[c] != 0;
        if constexpr (!kFullOnly) {
            const uint64_t bal = __ballot(smooth && q == 0);
            if ((threadIdx.x & 63) == 0 && bal) t.gated[lb * 4u + (threadIdx.x >> 6)] += (uint32_t)__popcll(bal);
        }
    } else if (kEval && t.gate) {
        const int lane = threadIdx.x & 63, gq = lane >> 4;
        const uint32_t cm = lb * 64u + (threadIdx.x >> 6) * 16u + (lane & 15);
        const bool vm = cm < N;
        const D4 Uo = ld4(U, vm ? cm : 0);
        const ind::CellRef r = ind::cell_ref_col(t.gamma, Uo.v, gq, lane & 15);
        // lane (column cell, group gq) holds members 4 kc + gq, kc = 0..2
        double d[3][4], smax[4] = {0.0, 0.0, 0.0, 0.0};
        bool ok = r.ok;
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) {
            const int m = 4 * kc + gq;
            if (m == 0) continue;
            const uint32_t pos = loc[teno_mem_off(lb, Mp, m - 1, cm & 63u, 0)];  // stencil 0
            const double2 u01 = sh_u[2 * pos], u23 = sh_u[2 * pos + 1];
            const double Um[4] = {u01.x, u01.y, u23.x, u23.y};
            ok = ind::member_diff(t.gamma, r, Um, d[kc]) && ok;
#pragma unroll
            for (int v = 0; v < 4; ++v) smax[v] = fmax(smax[v], fabs(d[kc][v]));
        }
        // stencil-wide maxima and validity: the 4 lanes of a column
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            smax[v] = fmax(smax[v], __shfl_xor(smax[v], 16));
            smax[v] = fmax(smax[v], __shfl_xor(smax[v], 32));
        }
        int okc = ok ? 1 : 0;
        okc &= __shfl_xor(okc, 16);
        okc &= __shfl_xor(okc, 32);
        // the 4 reciprocals of the maxima: lane gq takes variable gq's
        double is[4];
        const double is_mine = ind::inv_scale(sel4(smax, gq));
#pragma unroll
        for (int v = 0; v < 4; ++v) is[v] = __shfl(is_mine, (lane & 15) + 16 * v);
        ind::v4s xb[3];
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) {
            const int m = 4 * kc + gq;
            double x[4];
#pragma unroll
            for (int v = 0; v < 4; ++v)
                x[v] = !okc ? 0.0 : (m == 0 ? ind::scale_feature(smax[v]) : ind::pattern_feature(d[kc][v], is[v]));
            xb[kc] = ind::pack_feats(x);
        }
        const float z = ind::mlp_wave(ind::lds_net(t.net, sh_net), xb, lane);
        const int sm = (vm && okc && z >= t.net.logit_threshold) ? 1 : 0;
        if (vm && gq == 0) t.gate_dec[cm] = (uint8_t)sm;
        const uint64_t bal = __ballot(sm && gq == 0);
        if (lane == 0 && bal) t.gated[lb * 4u + (threadIdx.x >> 6)] += (uint32_t)__popcll(bal);
        if constexpr (GM == 3) {
            // mark the wave when a valid cell of it was not gated
            const uint64_t ung = __ballot(vm && !sm && gq == 0);
            if (lane == 0) t.blk_mixed[lb * 4u + (threadIdx.x >> 6)] = ung ? 1 : 0;
        }
        smooth = __shfl(sm, lane >> 2) != 0;
    }

    const uint32_t info0 = g.info[cc];
    const D4 uc = ld4(U, cc);
    const uint32_t cl = c & 63u;
    // Quad exchanges of the plain pass and of the gated passes that keep no
    // network in LDS (reuse, ungated cells), with 8-bit tiles, go through LDS: 4
    // planes of 258 doubles per block (the 2-double pad puts the 4 lanes of
    // a quad on distinct banks when they read one plane), 8 LDS instructions
    // per 4 x 4 transpose instead of the DPP form's selects and moves. A
    // wave reads only its own lanes' entries, and a wave's LDS accesses run
    // in order, so wave-scope fences (compiler ordering) suffice.
    constexpr bool kLdsX = (GM == 0 || GM == 4 || GM == 5) && TILE == 256;
    constexpr int kXs = 258;
    __shared__ double xbuf[kLdsX ? 4 * kXs : 1];
    auto wave_sync = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // out[j] = x[q] of quad lane j
    auto exchange = [&](const double* x, double* out) {
        if constexpr (kLdsX) {
            const uint32_t tid = threadIdx.x, base = tid & ~3u;
#pragma unroll
            for (int i = 0; i < 4; ++i) xbuf[i * kXs + tid] = x[i];
            wave_sync();
#pragma unroll
            for (int j = 0; j < 4; ++j) out[j] = xbuf[q * kXs + base + j];
            wave_sync();
        } else {
            quad_transpose(x, out, q);
        }
    };
    // out[j] = f[fq] of quad lane j, fq < FQ
    auto gather = [&](const double* f, double (*out)[4]) {
        if constexpr (kLdsX && FQ <= 4) {
            const uint32_t tid = threadIdx.x, base = tid & ~3u;
#pragma unroll
            for (int fq = 0; fq < FQ; ++fq) xbuf[fq * kXs + tid] = f[fq];
            wave_sync();
#pragma unroll
            for (int fq = 0; fq < FQ; ++fq)
#pragma unroll
                for (int j = 0; j < 4; ++j) out[fq][j] = xbuf[fq * kXs + base + j];
            wave_sync();
        } else {
#pragma unroll
            for (int fq = 0; fq < FQ; ++fq) quad_gather(f[fq], out[fq], q);
        }
    };
    // the member update a += A+_i (ta_i (u_i - u_c)) of member i of this
    // lane's stencil q; a wave's 64 lanes read 64 consecutive entries
    auto member = [&](int i, double (&a)[KP][4]) {
        const size_t o = teno_mem_off(lb, Mp, i, cl, q);
        const uint32_t pos = loc[o];
        const double2 u01 = sh_u[2 * pos], u23 = sh_u[2 * pos + 1];
        const D4 un = {{u01.x, u01.y, u23.x, u23.y}};
        double b[4];
        if constexpr (CT) {
#pragma unroll
            for (int v = 0; v < 4; ++v) b[v] = un.v[v] - uc.v[v];
        } else {
            const double tt = t.ta[o];
#pragma unroll
            for (int v = 0; v < 4; ++v) b[v] = tt * (un.v[v] - uc.v[v]);
        }
        const double* Ai = t.A + teno_a_off(lb, Mp, KP, i, 0, cl, q);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const double Ak = Ai[(size_t)k * 256];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[k][v] = a[k][v] + Ak * b[v];
        }
    };

    const double ucv = sel4(uc.v, q);  // lane-dependent pick by selects, not a stack array
    double f[FQ];
    bool central = false;
    // the central stencil's linear reconstruction through its precombined
    // operator (gated cells): quad lane q sums face points q and q + 4 for
    // all 4 variables over the members, a quad transpose hands lane v its
    // variable's face points
    auto linear_faces = [&]() {
        double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
        const bool has0 = q < FQ, has1 = q + 4 < FQ;
        if (valid && has0) {
            auto add = [&](int i) {
                const double* Lw = t.lin + teno_lin_off(lb, Mp, i, cl, FQ, q);
                const uint32_t pos = loc[teno_mem_off(lb, Mp, i, cl, 0)];  // central stencil
                const double2 u01 = sh_u[2 * pos], u23 = sh_u[2 * pos + 1];
                const double d[4] = {u01.x - uc.v[0], u01.y - uc.v[1], u23.x - uc.v[2], u23.y - uc.v[3]};
                const double w0 = Lw[0];
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[0][v] = acc[0][v] + w0 * d[v];
                if (has1) {
                    const double w1 = Lw[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[1][v] = acc[1][v] + w1 * d[v];
                }
            };
            if constexpr (MP > 0) {
#pragma unroll
                for (int i = 0; i < MP; ++i) add(i);
            } else {
                for (int i = 0; i < Mp; ++i) add(i);
            }
        }
        double y[4];
        exchange(acc[0], y);
#pragma unroll
        for (int fq = 0; fq < (FQ < 4 ? FQ : 4); ++fq) f[fq] = ucv + y[fq];
        if constexpr (FQ > 4) {
            exchange(acc[1], y);
#pragma unroll
            for (int fq = 4; fq < FQ; ++fq) f[fq] = ucv + y[fq - 4];
        }
        central = true;
    };

    // full TENO for this cell (the plain path)
    auto full_teno = [&]() {
        const bool active = valid && ((info0 >> (20 + q)) & 1u);

        // ---- phase A: lane q = stencil
        double a[KP][4];
#pragma unroll
        for (int k = 0; k < KP; ++k)
#pragma unroll
            for (int v = 0; v < 4; ++v) a[k][v] = 0.0;
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        if (active) {
            if constexpr (MP > 0) {
#pragma unroll
                for (int i = 0; i < MP; ++i) member(i, a);
            } else {
                for (int i = 0; i < Mp; ++i) member(i, a);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double si = 0.0;
#pragma unroll
                for (int k = 0; k < KP; ++k) {
                    double oia = 0.0;
#pragma unroll
                    for (int j = 0; j < KP; ++j) oia = oia + t.OI[k * KP + j] * a[j][v];
                    si = si + a[k][v] * oia;
                }
                w[v] = 1.0 / pow6_cr(si + 1.0e-12);
            }
        }

        // ---- exchange: lane q = variable
        double ws[4];
        double at[4][KP];
        exchange(w, ws);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            double x[4] = {a[k][0], a[k][1], a[k][2], a[k][3]}, y[4];
            exchange(x, y);
#pragma unroll
            for (int st = 0; st < 4; ++st) at[st][k] = y[st];
        }

        // ---- phase B: cut-off (face_reconstruction.cpp:948-981)
        constexpr int G = 4;
        const double CT1 = 1.0e-7, CT2 = 1.0e-5;
        double sd = 0.0;
#pragma unroll
        for (int st = 1; st < G; ++st) sd += ws[st];
        if (ws[0] / (sd + ws[0]) > CT1) {
            central = true;
            ws[0] = 1.0;
#pragma unroll
            for (int st = 1; st < G; ++st) ws[st] = 0.0;
        } else {
            if (t.variant != 0) ws[0] = 0.0;
#pragma unroll
            for (int st = 1; st < G; ++st) {
                if (ws[st] / sd > CT2) ws[st] = (1.0 / t.K);
                else if (t.variant != 0) ws[st] = 0.0;
            }
            double s2 = 0.0;
#pragma unroll
            for (int st = 1; st < G; ++st) s2 += ws[st];
#pragma unroll
            for (int st = 1; st < G; ++st) ws[st] /= s2;
        }
        double pw[G][KP];
#pragma unroll
        for (int st = 0; st < G; ++st)
#pragma unroll
            for (int k = 0; k < KP; ++k) pw[st][k] = ws[st] * at[st][k];

        // face points (face_reconstruction.cpp:985-1031)
        if constexpr (CT) {
            // keep the small table's loads here: hoisted to the kernel's
            // start (their addresses are known from info0) they would hold
            // 30 registers through the stencil loop
            __asm__ volatile("" ::: "memory");
        }
#pragma unroll
        for (int fq = 0; fq < FQ; ++fq) {
            double bk[KP];
            if constexpr (CT) {
                static_assert(FQ == 3, "compact tables: one face point per face");
                // the face's reference edge, from the slot-0 info word
                const double* B = t.fbc + ((info0 >> (24 + 2 * fq)) & 3u) * KP;
#pragma unroll
                for (int k = 0; k < KP; ++k) bk[k] = B[k];
            } else if constexpr (DV) {
                static_assert(FQ == 3 && KP == 5, "derived tables: p = 2, one face point per face");
                // face slot fq runs from cell node ia to node ib: the edge
                // (lo, hi) from bits 24 + 2 fq, its direction from the flip bit
                const double2* cv = reinterpret_cast<const double2*>(t.cvert) + 3 * (size_t)cc;
                const double2 n0 = cv[0], n1 = cv[1], n2 = cv[2];
                const TriFrame F = tri_frame(n0.x, n0.y, n1.x, n1.y, n2.x, n2.y);
                const uint32_t e = (info0 >> (24 + 2 * fq)) & 3u;
                const bool flip = ((info0 >> face_flip_shift(fq)) & 1u) != 0u;
                const uint32_t lo = e == 1u ? 1u : 0u, hi = e == 0u ? 1u : 2u;
                const uint32_t ia = flip ? hi : lo, ib = flip ? lo : hi;
                const double xa = ia == 0u ? n0.x : (ia == 1u ? n1.x : n2.x);
                const double ya = ia == 0u ? n0.y : (ia == 1u ? n1.y : n2.y);
                const double xb = ib == 0u ? n0.x : (ib == 1u ? n1.x : n2.x);
                const double yb = ib == 0u ? n0.y : (ib == 1u ? n1.y : n2.y);
                double ax, ay, bx, by;
                to_frame(F, xa, ya, ax, ay);
                to_frame(F, xb, yb, bx, by);
                const double s = (t.qf + 1.0) * 0.5;
                const double xq = s * (bx - ax) + ax, yq = s * (by - ay) + ay;
                // the reference weighting adds ipt_k over the cell's own area in its frame
                const double ta0 = t.variant != 0 ? 1.0 : frame_area(F, n0.x, n0.y, n1.x, n1.y, n2.x, n2.y);
#pragma unroll
                for (int k = 0; k < KP; ++k) {
                    const double psi =
                        legendre_le2(t.poly[2 * k + 2], xq) * legendre_le2(t.poly[2 * k + 3], yq);
                    bk[k] = t.variant != 0 ? psi - t.ipt[k + 1] : psi + (t.ipt[k + 1] / ta0);
                }
            } else {
                const double* B = t.fbasis + (((size_t)lb * FQ + fq) * KP) * 64 + (c & 63u);
#pragma unroll
                for (int k = 0; k < KP; ++k) bk[k] = B[k * 64];
            }
            double acc = ucv;
#pragma unroll
            for (int st = 0; st < G; ++st) {
                if (ws[st] == 0.0) continue;
                acc = acc + (ws[st] * 0.0) * t.basis0;
#pragma unroll
                for (int k = 0; k < KP; ++k) acc = acc + pw[st][k] * bk[k];
            }
            f[fq] = acc;
        }
    };
    // a cell another pass of the split reconstructs: the cell average
    // stands in (positivity leaves it alone) and nothing is stored
    auto skip = [&]() {
#pragma unroll
        for (int fq = 0; fq < FQ; ++fq) f[fq] = ucv;
    };

    if constexpr (kFullOnly) {
        if (smooth) skip();
        else full_teno();
    } else if (GM != 0 && t.gate && __all(smooth || !valid)) {
        // every cell of the wave takes its central stencil
        smooth = true;  // lanes past the end too: they store nothing
        linear_faces();
    } else if (GM != 0 && smooth) {
        linear_faces();  // a gated cell in a wave with ungated ones
    } else if constexpr (kLinOnly) {
        skip();
    } else {
        full_teno();
    }
    const int v = q;

    if (t.variant != 0) {
        // positivity scaling of the cell's face states (Zhang & Shu 2010),
        // computed redundantly by the 4 lanes, in the oracle's order
        double uf[FQ][4];
        gather(f, uf);
        const double Uc[4] = {uc.v[0], uc.v[1], uc.v[2], uc.v[3]};
        const double pc = pressure_raw(ph.gamma, Uc);
        const double eps = fmin(1.0e-13, fmin(Uc[0], pc));
        double theta = 1.0;
#pragma unroll
        for (int fq = 0; fq < FQ; ++fq) {
            const double* u = uf[fq];
            double wv[4];
            double tt = 1.0;
            if (u[0] < eps) tt = (Uc[0] - eps) / (Uc[0] - u[0]);
#pragma unroll
            for (int k = 0; k < 4; ++k) wv[k] = Uc[k] + tt * (u[k] - Uc[k]);
            if (pressure_raw(ph.gamma, wv) < eps) {
                double lo = 0.0, hi = tt;
                for (int it = 0; it < 40; it++) {
                    const double mid = 0.5 * (lo + hi);
#pragma unroll
                    for (int k = 0; k < 4; ++k) wv[k] = Uc[k] + mid * (u[k] - Uc[k]);
                    if (pressure_raw(ph.gamma, wv) < eps) hi = mid;
                    else lo = mid;
                }
                tt = lo;
            }
            theta = fmin(theta, tt);
        }
        if (theta < 1.0)
#pragma unroll
            for (int fq = 0; fq < FQ; ++fq) f[fq] = ucv + theta * (f[fq] - ucv);
    }
    const uint64_t cmask = __ballot(central);
    if (valid && (kLinOnly ? smooth : kFullOnly ? !smooth : true)) {
#pragma unroll
        for (int fq = 0; fq < FQ; ++fq) fs[((size_t)fq * N + c) * 4 + v] = f[fq];
        if (dec && v == 0) dec[c] = (uint8_t)((cmask >> ((threadIdx.x & 63u) & ~3u)) & 0xFu);
    }
}

// OCC: waves per SIMD the register allocation targets (0: the compiler's
// own choice); an autotuner picks it per mesh, results are identical
template <int KP, int MP, int FQ, typename IdxT, int TILE, int GM, int OCC, bool CT = false, bool DV = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC == 0 ? 1 : OCC, OCC == 0 ? 8 : OCC)))
void teno_kernel(Geom g, Teno t, const double* __restrict__ U,
                                                   double* __restrict__ fs, uint8_t* __restrict__ dec,
                                                   Physics ph, uint32_t blocks_per_xcd) {
    // dynamic LDS: the tile's states (TILE cells), then the network when gated
    extern __shared__ __align__(16) unsigned char teno_smem[];
    const uint32_t lb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    teno_block<KP, MP, FQ, IdxT, TILE, GM, CT, DV>(g, t, U, fs, dec, ph, lb, reinterpret_cast<double2*>(teno_smem),
                                                   teno_smem + TILE * 32);
}

// case of global block gb: the last case whose first block is <= gb
__device__ inline int case_of_block(const CaseBufs* __restrict__ cs, int n, uint32_t gb, bool flux) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((flux ? cs[mid].blk_flux0 : cs[mid].blk_teno0) <= gb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline const double* case_buf(const CaseBufs& c, int role) {
    return role == kBufUn ? c.Un : (role == kBufUa ? c.Ua : c.Ub);
}
__device__ inline double* case_buf_w(const CaseBufs& c, int role) {
    return role == kBufUn ? c.Un : (role == kBufUa ? c.Ua : (role == kBufUb ? c.Ub : nullptr));
}

// The TENO pass of every case of an ensemble in one launch: the grid covers
// all cases' 64-cell blocks (XCD-aware order over the whole grid), each
// block finds its case in the offset table and runs that case's block.
template <int KP, int MP, int FQ, typename IdxT, int TILE, int GM, int OCC, bool DV = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC == 0 ? 1 : OCC, OCC == 0 ? 8 : OCC)))
void teno_batch_kernel(const CaseBufs* __restrict__ cases, int n_cases, int in_role,
                                                         uint32_t total, Physics ph, uint32_t blocks_per_xcd) {
    extern __shared__ __align__(16) unsigned char teno_smem[];
    const uint32_t gb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    if (gb >= total) return;
    const CaseBufs& cs = cases[case_of_block(cases, n_cases, gb, false)];
    teno_block<KP, MP, FQ, IdxT, TILE, GM, false, DV>(cs.g, cs.t, case_buf(cs, in_role), cs.fs, cs.dec, ph,
                                                      gb - cs.blk_teno0, reinterpret_cast<double2*>(teno_smem),
                                                      teno_smem + TILE * 32);
}

// ------------------------------------------------------------ viscous terms
// Navier-Stokes extension (no reference counterpart). Sutherland viscosity.
__device__ inline double viscosity(const Physics& ph, double T) {
    if (ph.S_mu < 0.0) return ph.mu_ref;
    const double t = T / ph.T_mu;
    return ph.mu_ref * (t * sqrt(t)) * ((ph.T_mu + ph.S_mu) / (T + ph.S_mu));
}

// Values of (u, v, T) a boundary face holds for the gradients: no-slip wall
// (u, v) = 0; slip wall and symmetry the tangential velocity; fixed (u, p, T)
// its values; extrapolation and p_out the cell's own (no contribution).
__device__ inline void boundary_uvT(const BCParam& b, double n0, double n1, const double* c, double* f) {
    if (b.type == 5) {
        f[0] = 0.0; f[1] = 0.0; f[2] = c[2];
    } else if (b.type == 0 || b.type == 2) {
        const double un = c[0] * n0 + c[1] * n1;
        f[0] = c[0] - un * n0; f[1] = c[1] - un * n1; f[2] = c[2];
    } else if (b.type == 3) {
        f[0] = b.u0; f[1] = b.u1; f[2] = b.T;
    } else {
        f[0] = c[0]; f[1] = c[1]; f[2] = c[2];
    }
}

// One thread per cell: (u, v, T, mu) of the state and the weighted
// least-squares gradients of u, v and T over the face neighbours (weights
// 1/|d|^2, d from the centroid to the neighbour's centroid or to the boundary
// face's midpoint), exact for linear fields.
__global__ __launch_bounds__(256) void viscous_grad_kernel(Geom g, const double* __restrict__ U, Physics ph,
                                                           double* __restrict__ vp, double* __restrict__ vg,
                                                           uint32_t blocks_per_xcd) {
    const uint32_t lb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    const uint32_t c = lb * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    double Wc[5];
    const D4 uc = ld4(U, c);
    prims(ph, uc.v, Wc);
    const double pc[3] = {Wc[0], Wc[1], Wc[3]};
    double a11 = 0.0, a12 = 0.0, a22 = 0.0, b[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (uint32_t l = 0; l < g.S; ++l) {
        const size_t o = (size_t)l * g.N + c;
        const uint32_t inf = g.info[o];
        const double dx = g.dcx[o], dy = g.dcy[o];
        double pf[3];
        if (!(inf & kBoundaryBit)) {
            double Wn[5];
            const D4 un = ld4(U, (uint32_t)g.nb[o]);
            prims(ph, un.v, Wn);
            pf[0] = Wn[0]; pf[1] = Wn[1]; pf[2] = Wn[3];
        } else {
            boundary_uvT(g.bc[(inf >> 4) & 0xFu], g.nx[o], g.ny[o], pc, pf);
        }
        const double w = 1.0 / (dx * dx + dy * dy);
        a11 += w * dx * dx;
        a12 += w * dx * dy;
        a22 += w * dy * dy;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double dp = w * (pf[k] - pc[k]);
            b[k][0] += dx * dp;
            b[k][1] += dy * dp;
        }
    }
    const double id = 1.0 / (a11 * a22 - a12 * a12);
    double gr[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gr[2 * k] = (a22 * b[k][0] - a12 * b[k][1]) * id;
        gr[2 * k + 1] = (a11 * b[k][1] - a12 * b[k][0]) * id;
    }
    st4(vp, c, D4{{pc[0], pc[1], pc[2], viscosity(ph, pc[2])}});
    double2* G2 = reinterpret_cast<double2*>(vg + 6 * (size_t)c);
    G2[0] = make_double2(gr[0], gr[1]);
    G2[1] = make_double2(gr[2], gr[3]);
    G2[2] = make_double2(gr[4], gr[5]);
}

// Viscous flux through a face, in the face's orientation (unit normal n from
// its side-0 cell L to R): G = (0, tau n, (u tau + k grad T) n). Face
// gradients are the mean of the two cells' gradients with the directional
// derivative along d = x_R - x_L replaced by (phi_R - phi_L) / |d|; the face
// velocity is the mean of both cells' linear extrapolations to the face
// midpoint. On a boundary face the cell's gradient is corrected toward the
// boundary value along the centroid-to-midpoint vector. All of it is exact
// for linear fields, and both cells of a face evaluate the same operations on
// the same values, so they agree bit for bit.
__device__ inline void viscous_face(const Physics& ph, const Geom& g, const StageArgs& a, size_t o, uint32_t c,
                                    uint32_t inf, bool bnd, uint32_t side, double n0, double n1, double* G) {
    G[0] = 0.0; G[1] = 0.0; G[2] = 0.0; G[3] = 0.0;
    int bt = -1;
    if (bnd) {
        bt = g.bc[(inf >> 4) & 0xFu].type;
        if (bt == 0 || bt == 2) return;  // symmetry, slip wall: no viscous stress or heat flux
    }
    const uint32_t L = (!bnd && side) ? (uint32_t)g.nb[o] : c;
    const D4 pL = ld4(a.vp, L);
    const double2* gL2 = reinterpret_cast<const double2*>(a.vg + 6 * (size_t)L);
    const double2 gL0 = gL2[0], gL1 = gL2[1], gL3 = gL2[2];
    double dx = g.dcx[o], dy = g.dcy[o];
    if (side) { dx = -dx; dy = -dy; }
    const double il = 1.0 / sqrt(dx * dx + dy * dy);
    const double tx = dx * il, ty = dy * il;
    double gx[3], gy[3], uf, vf, mu;
    const double phL[3] = {pL.v[0], pL.v[1], pL.v[2]};
    if (!bnd) {
        const uint32_t R = side ? c : (uint32_t)g.nb[o];
        const D4 pR = ld4(a.vp, R);
        const double2* gR2 = reinterpret_cast<const double2*>(a.vg + 6 * (size_t)R);
        const double2 gR0 = gR2[0], gR1 = gR2[1], gR3 = gR2[2];
        gx[0] = 0.5 * (gL0.x + gR0.x); gy[0] = 0.5 * (gL0.y + gR0.y);
        gx[1] = 0.5 * (gL1.x + gR1.x); gy[1] = 0.5 * (gL1.y + gR1.y);
        gx[2] = 0.5 * (gL3.x + gR3.x); gy[2] = 0.5 * (gL3.y + gR3.y);
        const double phR[3] = {pR.v[0], pR.v[1], pR.v[2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double corr = (phR[k] - phL[k]) * il - (gx[k] * tx + gy[k] * ty);
            gx[k] += corr * tx;
            gy[k] += corr * ty;
        }
        // centroid-to-midpoint vectors of L and R (own slot and the neighbour's slot)
        const size_t on = (size_t)(inf & 3u) * g.N + (uint32_t)g.nb[o];
        const size_t oL = side ? on : o, oR = side ? o : on;
        const double mLx = g.mfx[oL], mLy = g.mfy[oL], mRx = g.mfx[oR], mRy = g.mfy[oR];
        uf = 0.5 * ((pL.v[0] + (gL0.x * mLx + gL0.y * mLy)) + (pR.v[0] + (gR0.x * mRx + gR0.y * mRy)));
        vf = 0.5 * ((pL.v[1] + (gL1.x * mLx + gL1.y * mLy)) + (pR.v[1] + (gR1.x * mRx + gR1.y * mRy)));
        mu = 0.5 * (pL.v[3] + pR.v[3]);
    } else {
        gx[0] = gL0.x; gy[0] = gL0.y;
        gx[1] = gL1.x; gy[1] = gL1.y;
        gx[2] = gL3.x; gy[2] = gL3.y;
        double pf[3];
        boundary_uvT(g.bc[(inf >> 4) & 0xFu], n0, n1, phL, pf);
        if (bt != 1 && bt != 4) {  // a prescribed value: correct toward it
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double corr = (pf[k] - phL[k]) * il - (gx[k] * tx + gy[k] * ty);
                gx[k] += corr * tx;
                gy[k] += corr * ty;
            }
        }
        uf = pf[0];
        vf = pf[1];
        mu = pL.v[3];
    }
    const double div = gx[0] + gy[1];
    const double txx = mu * (2.0 * gx[0] - (2.0 / 3.0) * div);
    const double tyy = mu * (2.0 * gy[1] - (2.0 / 3.0) * div);
    const double txy = mu * (gy[0] + gx[1]);
    G[1] = txx * n0 + txy * n1;
    G[2] = txy * n0 + tyy * n1;
    // adiabatic no-slip wall: u = 0 and no heat flux
    if (bt == 5) return;
    const double k = mu * ph.cp / ph.Pr;
    G[3] = uf * G[1] + vf * G[2] + k * (gx[2] * n0 + gy[2] * n1);
}

// ------------------------------------------------------------ flux + update
// One thread per cell: the residual of the cell in the canonical face order,
// divided by the volume, then the stage's SSP-RK update (no atomics).
// RS: Riemann solver (compile-time, so each instantiation carries one
// solver's code and registers). Blocks follow the XCD-aware order of the
// TENO kernel, so the neighbour face states a block gathers were written by
// the same XCD.
// PARTS: face sums taken, bit 0 interior faces, bit 1 boundary faces (3: the
// whole residual; 1 or 2 only with kRhsOnly).
// PP: the positivity-preserving flux limiter (pp_limit) on every face, for
// the stage updates (not the residual-only operator, which has no dt).
template <int RS, int PARTS = 3, bool PP = false>
__device__ __forceinline__ void flux_cell(const Geom& g, const StageArgs& a, const Physics& ph, uint32_t c) {
    const uint32_t N = g.N;
    if (c >= N) return;
    const uint32_t Q = g.Q;
    const uint32_t info0 = g.info[c];
    const uint32_t code = (info0 >> 8) & 0xFFu, nact = (info0 >> 16) & 0x7u;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    D4 um{};
    double vc = 0.0, pdt = 0.0;
    if constexpr (PP) {
        um = ld4(a.Uin, c);
        vc = g.volume[c];
        pdt = *a.dt;
    }
    for (uint32_t rk = 0; rk < nact; ++rk) {
        const uint32_t l = (code >> (2 * rk)) & 3u;
        const size_t o = (size_t)l * N + c;
        const uint32_t inf = g.info[o];
        const bool bnd = (inf & kBoundaryBit) != 0;
        if constexpr (PARTS != 3) {
            if (bnd ? !(PARTS & 2) : !(PARTS & 1)) continue;
        }
        const uint32_t side = (inf >> 2) & 1u;
        const double n0 = g.nx[o], n1 = g.ny[o], ar = g.area[o];
        double F[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t q = 0; q < Q; ++q) {
            D4 own, oth;
            if (a.fo) {
                own = ld4(a.Uin, c);
            } else {
                own = ld4(a.fs, (size_t)(l * Q + q) * N + c);
            }
            double Ft[4];
            if (!bnd) {
                const uint32_t nbc = (uint32_t)g.nb[o];
                const uint32_t nbs = inf & 3u;
                oth = a.fo ? ld4(a.Uin, nbc) : ld4(a.fs, (size_t)(nbs * Q + q) * N + nbc);
                double L[4], R[4];  // selected by value (a selected pointer can put them in scratch)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    L[i] = side ? oth.v[i] : own.v[i];
                    R[i] = side ? own.v[i] : oth.v[i];
                }
                point_flux(ph, g.bc, RS, a.corrected, false, 0, n0, n1, L, R, Ft);
            } else {
                // the ghost state an extrapolation boundary copies: the face
                // state itself (boundary_extrapolation.cpp:38-55), or in the
                // corrected variant the cell mean
                const D4 ghost = a.corrected ? ld4(a.Uin, c) : own;
                point_flux(ph, g.bc, RS, a.corrected, true, (inf >> 4) & 0xFu, n0, n1, own.v, ghost.v, Ft);
            }
            const double wq = g.qw[q];
#pragma unroll
            for (int i = 0; i < 4; i++) F[i] += wq * Ft[i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) F[i] *= 0.5;
        if constexpr (PP) {
            // in side order (side 0 first), so both cells of a face agree bit for bit
            const uint32_t nbc = bnd ? c : (uint32_t)g.nb[o];
            const D4 mo = bnd ? um : ld4(a.Uin, nbc);
            const double vo = bnd ? vc : g.volume[nbc];
            double U0[4], U1[4];  // selected by value: a selected pointer would put the means in scratch
#pragma unroll
            for (int i = 0; i < 4; i++) {
                U0[i] = side ? mo.v[i] : um.v[i];
                U1[i] = side ? um.v[i] : mo.v[i];
            }
            const double V0 = side ? vo : vc, V1 = side ? vc : vo;
            const double Sd = (double)g.S;
            pp_limit(ph, g.bc, bnd, (inf >> 4) & 0xFu, n0, n1, F, U0, Sd * pdt * ar / V0, U1, Sd * pdt * ar / V1);
        }
        if (ph.viscous) {
            double G[4];
            viscous_face(ph, g, a, o, c, inf, bnd, side, n0, n1, G);
#pragma unroll
            for (int i = 1; i < 4; i++) F[i] -= G[i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] += side ? ar * F[i] : -ar * F[i];
    }
    const double vol = g.volume[c];
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] /= vol;

    if (a.op == kRhsOnly) {
        st4(a.rhs, c, D4{{r[0], r[1], r[2], r[3]}});
        return;
    }
    const double dt = *a.dt;
    const D4 un = ld4(a.Un, c);
    D4 out, wa;
    bool final_stage = false;
    switch (a.op) {
    case kFE:
        for (int i = 0; i < 4; i++) out.v[i] = un.v[i] + dt * r[i];
        st4(a.Uout, c, out);
        final_stage = true;
        break;
    case kSSP1: {
        const double c0 = dt * (1.0 / 6.0);
        for (int i = 0; i < 4; i++) { out.v[i] = un.v[i] + dt * r[i]; wa.v[i] = un.v[i] + c0 * r[i]; }
        st4(a.Uout, c, out);
        st4(a.W, c, wa);
        break;
    }
    case kSSP2: {
        const D4 u1 = ld4(a.Uin, c);
        const D4 w0 = ld4(a.W, c);
        const double c1 = dt * (1.0 / 6.0), d4 = dt / 4.0;
        for (int i = 0; i < 4; i++) {
            const double y = 3.0 / 4.0 * un.v[i] + 1.0 / 4.0 * u1.v[i];
            out.v[i] = y + d4 * r[i];
            wa.v[i] = w0.v[i] + c1 * r[i];
        }
        st4(a.Uout, c, out);
        st4(a.W, c, wa);
        break;
    }
    case kSSP3: {
        const D4 w0 = ld4(a.W, c);
        const double c2 = dt * (2.0 / 3.0);
        for (int i = 0; i < 4; i++) out.v[i] = w0.v[i] + c2 * r[i];
        st4(a.Uout, c, out);
        final_stage = true;
        break;
    }
    case kRK4_1: case kRK4_2: case kRK4_3: {
        const double half = dt / 2.0;
        const double cs = a.op == kRK4_1 ? dt * (1.0 / 6.0) : dt * (1.0 / 3.0);
        const double step = a.op == kRK4_3 ? dt : half;
        D4 w0 = un;
        if (a.op != kRK4_1) w0 = ld4(a.W, c);
        for (int i = 0; i < 4; i++) { out.v[i] = un.v[i] + step * r[i]; wa.v[i] = w0.v[i] + cs * r[i]; }
        st4(a.Uout, c, out);
        st4(a.W, c, wa);
        break;
    }
    case kRK4_4: {
        const D4 w0 = ld4(a.W, c);
        const double c3 = dt * (1.0 / 6.0);
        for (int i = 0; i < 4; i++) out.v[i] = w0.v[i] + c3 * r[i];
        st4(a.Uout, c, out);
        final_stage = true;
        break;
    }
    default: break;
    }
    if (final_stage && a.prims) {
        double W5[5];
        prims(ph, out.v, W5);
        st4(a.prims, c, D4{{W5[0], W5[1], W5[2], W5[3]}});
        a.h[c] = W5[4];
    }
}

template <int RS, bool PP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PP ? 4 : 1, 8)))
void flux_update_kernel(Geom g, StageArgs a, Physics ph, uint32_t blocks_per_xcd) {
    const uint32_t lb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    flux_cell<RS, 3, PP>(g, a, ph, lb * blockDim.x + threadIdx.x);
}

// The residual restricted to the interior faces (PARTS = 1, the
// reference's calc_rhs_interior) or to the boundary faces (PARTS = 2, its
// calc_rhs_boundaries, i.e. every Boundary::apply), divided by the volume.
template <int RS, int PARTS>
__global__ __launch_bounds__(256) void rhs_parts_kernel(Geom g, StageArgs a, Physics ph, uint32_t blocks_per_xcd) {
    const uint32_t lb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    flux_cell<RS, PARTS>(g, a, ph, lb * blockDim.x + threadIdx.x);
}

// The flux + stage update of every case in one launch (blocks of blockDim.x
// cells; the case table's block offsets count blocks of that size).
template <int RS, bool PP>
__global__ __launch_bounds__(256) void flux_batch_kernel(const CaseBufs* __restrict__ cases, int n_cases,
                                                         BatchStage st, uint32_t total, Physics ph,
                                                         uint32_t blocks_per_xcd) {
    const uint32_t gb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    if (gb >= total) return;
    const CaseBufs& cs = cases[case_of_block(cases, n_cases, gb, true)];
    StageArgs a{};
    a.op = st.op;
    a.riemann = RS;
    a.fo = st.fo;
    a.corrected = st.corrected;
    a.Uin = case_buf(cs, st.in);
    a.Un = cs.Un;
    a.Uout = case_buf_w(cs, st.out);
    a.W = cs.W;
    a.prims = cs.prims;
    a.h = cs.h;
    a.rhs = cs.rhs;
    a.fs = cs.fs;
    a.dt = cs.dt;
    flux_cell<RS, 3, PP>(cs.g, a, ph, (gb - cs.blk_flux0) * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------ dt
// solver.cpp:592-706: per-cell spectral radius over the cell's faces in
// local order; the block's maximum, joined over the blocks by
// finalize_dt_kernel.
__device__ __forceinline__ void spectral_block(const Geom& g, const double* __restrict__ Un,
                                               const double* __restrict__ pr, const Physics& ph,
                                               double* __restrict__ lam_cell, unsigned long long* __restrict__ lam_blk,
                                               uint32_t lb) {
    const uint32_t N = g.N;
    const uint32_t c = lb * blockDim.x + threadIdx.x;
    double lam = 0.0;
    bool ok = false;
    if (c < N) {
        const D4 wc = ld4(pr, c);
        const double rc = Un[4 * (size_t)c];
        const double sc = sqrt(ph.gamma * wc.v[2] / rc);
        double conv = 0.0, acou = 0.0;
        const uint32_t S = g.S;
        for (uint32_t l = 0; l < S; ++l) {
            const size_t o = (size_t)l * N + c;
            const uint32_t nbc = (uint32_t)g.nb[o];
            const double n0 = g.nx[o], n1 = g.ny[o], dxn = g.dxn[o];
            const D4 wn = ld4(pr, nbc);
            const double rn = Un[4 * (size_t)nbc];
            const double sn = sqrt(ph.gamma * wn.v[2] / rn);
            const double sf = 0.5 * (sc + sn);
            const double uf0 = 0.5 * (wc.v[0] + wn.v[0]), uf1 = 0.5 * (wc.v[1] + wn.v[1]);
            const double un = fabs(uf0 * n0 + uf1 * n1);
            conv += un / dxn;
            const double qq = sf / dxn;
            acou += qq * qq;
        }
        const double gf = 3.0 / S;
        conv *= 1.37 * gf;
        acou = 1.37 * sqrt(gf * acou);
        lam = conv + acou;
        if (ph.viscous) {
            // viscous radius max(4/3, gamma/Pr) (mu / rho) sum_f (A_f / V)^2
            const double V = g.volume[c];
            double s2 = 0.0;
            for (uint32_t l = 0; l < S; ++l) {
                const double q = g.area[(size_t)l * N + c] / V;
                s2 += q * q;
            }
            lam += fmax(4.0 / 3.0, ph.gamma / ph.Pr) * (viscosity(ph, wc.v[3]) / rc) * s2;
        }
        lam_cell[c] = lam;
        ok = lam == lam;  // NaN radii are ignored by the reference's max join
    }
    // non-negative doubles order like their bit patterns
    unsigned long long bits = ok ? (unsigned long long)__double_as_longlong(lam) : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o2 = __shfl_xor(bits, off);
        bits = o2 > bits ? o2 : bits;
    }
    __shared__ unsigned long long sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = sh[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = sh[i] > m ? sh[i] : m;
        lam_blk[lb] = m;  // the block's maximum; the finalize kernel joins the blocks
    }
}

// max over n per-block maxima (bit patterns of non-negative doubles) by one
// block; every thread returns the result
__device__ inline unsigned long long join_blocks(const unsigned long long* __restrict__ lam_blk, uint32_t n) {
    unsigned long long m = 0ull;
    const uint32_t st = blockDim.x;
    uint32_t i = threadIdx.x;
    // 8 loads in flight per thread (the max is order-free)
    for (; i + 7 * st < n; i += 8 * st) {
        unsigned long long v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = lam_blk[i + k * st];
#pragma unroll
        for (int k = 0; k < 8; ++k) m = v[k] > m ? v[k] : m;
    }
    for (; i < n; i += st) m = lam_blk[i] > m ? lam_blk[i] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o2 = __shfl_xor(m, off);
        m = o2 > m ? o2 : m;
    }
    __shared__ unsigned long long sh[16];
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    m = sh[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = sh[i] > m ? sh[i] : m;
    return m;
}

__global__ __launch_bounds__(256) void spectral_kernel(Geom g, const double* __restrict__ Un,
                                                       const double* __restrict__ pr, Physics ph,
                                                       double* __restrict__ lam_cell,
                                                       unsigned long long* __restrict__ lam_blk,
                                                       uint32_t blocks_per_xcd) {
    // each XCD walks one contiguous range of blocks (its neighbours' states stay in its L2)
    const uint32_t lb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    if (lb * blockDim.x >= g.N) return;  // past the last block (uniform)
    spectral_block(g, Un, pr, ph, lam_cell, lam_blk, lb);
}

// spectral radii of every case in one launch (each block inside one case, so
// the block max goes to that case's maximum)
__global__ __launch_bounds__(256) void spectral_batch_kernel(const CaseBufs* __restrict__ cases, int n_cases,
                                                             uint32_t total, Physics ph, uint32_t blocks_per_xcd) {
    const uint32_t gb = (blockIdx.x & 7u) * blocks_per_xcd + (blockIdx.x >> 3);
    if (gb >= total) return;
    const CaseBufs& cs = cases[case_of_block(cases, n_cases, gb, true)];
    spectral_block(cs.g, cs.Un, cs.prims, ph, cs.lam_cell, cs.lam_blk, gb - cs.blk_flux0);
}

// dt = cfl / max lambda over the per-block maxima (one block of 1024 threads)
// or the fixed dt; its history entry and the step counter
__global__ __launch_bounds__(1024) void finalize_dt_kernel(const unsigned long long* __restrict__ lam_blk,
                                                           uint32_t n_blk, double cfl, double fixed_dt, double* dt,
                                                           double* dt_hist, int64_t hist_cap, int64_t* step) {
    const unsigned long long m = fixed_dt > 0.0 ? 0ull : join_blocks(lam_blk, n_blk);
    if (threadIdx.x != 0) return;
    const double d = fixed_dt > 0.0 ? fixed_dt : cfl / __longlong_as_double((long long)m);
    *dt = d;
    if (dt_hist && step) dt_hist[*step % hist_cap] = d;
    if (step) *step += 1;
}

// one case's dt from its per-block maxima (blocks of bs cells), its history
// entry and step counter; every thread of the block takes part
__device__ inline void finalize_case(const CaseBufs& cs, double cfl, double fixed_dt, int64_t hist_cap, uint32_t bs) {
    const unsigned long long m = fixed_dt > 0.0 ? 0ull : join_blocks(cs.lam_blk, (cs.g.N + bs - 1) / bs);
    if (threadIdx.x != 0) return;
    const double d = fixed_dt > 0.0 ? fixed_dt : cfl / __longlong_as_double((long long)m);
    *cs.dt = d;
    cs.dt_hist[*cs.step % hist_cap] = d;
    *cs.step += 1;
}

// per-case dt of an ensemble: block k finalises case k
__global__ __launch_bounds__(256) void finalize_batch_kernel(const CaseBufs* __restrict__ cases, int n_cases,
                                                             double cfl, double fixed_dt, int64_t hist_cap,
                                                             uint32_t bs) {
    const int k = blockIdx.x;
    if (k >= n_cases) return;
    finalize_case(cases[k], cfl, fixed_dt, hist_cap, bs);
}

// ------------------------------------------------------------ persistent ensemble step
// The whole step of every case in one cooperative launch: the spectral radii
// and each case's dt, then per stage the TENO face states and the flux +
// stage update, with a grid-wide barrier between phases. The grid holds as
// many blocks as fit on the GPU at once; each phase's work items (64-cell
// TENO blocks, 256-cell flux / spectral blocks, over all cases through the
// offset table) are dealt out so that each XCD walks one contiguous eighth,
// as the per-phase kernels do. Every item runs the per-phase kernels' own
// per-cell code, so results are bit-identical to them.
// Grid-wide barrier of the persistent launch (every block resident: the
// cooperative launch guarantees it). Each wave first waits for its own
// stores to reach its XCD's L2; then one thread per block releases them past
// that L2 (each XCD has its own), arrives on a counter and waits for the
// generation the last arrival advances, and drops stale lines of its L2 once
// through. The waits poll with relaxed loads and no invalidation, so blocks
// that arrive early leave the L2 of those still working alone. A wait longer
// than spin_ticks of the device wall clock gives up and raises bar->err (the
// launch then still ends; the host reports the error and resets the state).
__device__ inline void grid_barrier(StepBarrier* bar, uint32_t n_blocks, long long spin_ticks) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t g = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == n_blocks - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every arrival's release is ordered before
            __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar->gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const long long t0 = wall_clock64();
            while (__hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(8);
                if (wall_clock64() - t0 > spin_ticks) {
                    __hip_atomic_store(&bar->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

template <int RS, bool TENO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void ensemble_step_kernel(const CaseBufs* __restrict__ cases, int n_cases,
                                                            uint32_t teno_total, uint32_t flux_total,
                                                            PersistentStep ps, Physics ph) {
    extern __shared__ __align__(16) unsigned char step_smem[];
    const uint32_t G = gridDim.x, xcd = blockIdx.x & 7u, rank = blockIdx.x >> 3;
    const uint32_t on_xcd = (G + 7u - xcd) >> 3;
    auto stamp = [&](int i) {
        if (ps.clocks && blockIdx.x == 0 && threadIdx.x == 0) ps.clocks[i] = wall_clock64();
    };
    stamp(0);
    auto each = [&](uint32_t total, auto&& f) {
        const uint32_t per = (total + 7u) >> 3, lo = xcd * per, hi = lo + per < total ? lo + per : total;
        for (uint32_t w = lo + rank; w < hi; w += on_xcd) {
            f(w);
            __syncthreads();  // the next item reuses the block's LDS
        }
    };
    if (ps.use_cfl) {
        each(flux_total, [&](uint32_t gb) {
            const CaseBufs& cs = cases[case_of_block(cases, n_cases, gb, true)];
            spectral_block(cs.g, cs.Un, cs.prims, ph, cs.lam_cell, cs.lam_blk, gb - cs.blk_flux0);
        });
        grid_barrier(ps.bar, G, ps.spin_ticks);
    }
    stamp(1);
    for (int k = (int)blockIdx.x; k < n_cases; k += (int)G) {
        finalize_case(cases[k], ps.cfl, ps.fixed_dt, ps.hist_cap, 256u);
        __syncthreads();
    }
    grid_barrier(ps.bar, G, ps.spin_ticks);
    stamp(2);
#pragma unroll 1
    for (int st = 0; st < ps.n_stages; ++st) {
        if constexpr (TENO) {
            each(teno_total, [&](uint32_t gb) {
                const CaseBufs& cs = cases[case_of_block(cases, n_cases, gb, false)];
                teno_block<5, 11, 3, uint8_t, 256, 0>(cs.g, cs.t, case_buf(cs, ps.in[st]), cs.fs, cs.dec, ph,
                                                      gb - cs.blk_teno0, reinterpret_cast<double2*>(step_smem),
                                                      step_smem + 256 * 32);
            });
            grid_barrier(ps.bar, G, ps.spin_ticks);
        }
        stamp(3 + 2 * st);
        each(flux_total, [&](uint32_t gb) {
            const CaseBufs& cs = cases[case_of_block(cases, n_cases, gb, true)];
            StageArgs a{};
            a.op = ps.op[st];
            a.riemann = RS;
            a.fo = ps.fo;
            a.corrected = ps.corrected;
            a.Uin = case_buf(cs, ps.in[st]);
            a.Un = cs.Un;
            a.Uout = case_buf_w(cs, ps.out[st]);
            a.W = cs.W;
            a.prims = cs.prims;
            a.h = cs.h;
            a.rhs = cs.rhs;
            a.fs = cs.fs;
            a.dt = cs.dt;
            flux_cell<RS>(cs.g, a, ph, (gb - cs.blk_flux0) * blockDim.x + threadIdx.x);
        });
        if (st + 1 < ps.n_stages || ps.clocks) grid_barrier(ps.bar, G, ps.spin_ticks);
        stamp(4 + 2 * st);
    }
}

__global__ void copy_kernel(const double* __restrict__ s, double* __restrict__ d, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// Streaming probes: the bandwidth HBM delivers to a plain access pattern
// (16 B per lane, grid-stride, 4 loads in flight per lane), the practical
// ceiling the gather kernels are compared with.
__global__ __launch_bounds__(256) void stream_copy_kernel(const double2* __restrict__ s, double2* __restrict__ d,
                                                          size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 a = s[i], b = s[i + stride], c = s[i + 2 * stride], e = s[i + 3 * stride];
        d[i] = a;
        d[i + stride] = b;
        d[i + 2 * stride] = c;
        d[i + 3 * stride] = e;
    }
    for (; i < n; i += stride) d[i] = s[i];
}

__global__ __launch_bounds__(256) void stream_read_kernel(const double2* __restrict__ s, double* __restrict__ out,
                                                          size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc0 = 0.0, acc1 = 0.0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 a = s[i], b = s[i + stride], c = s[i + 2 * stride], e = s[i + 3 * stride];
        acc0 += (a.x + b.x) + (c.x + e.x);
        acc1 += (a.y + b.y) + (c.y + e.y);
    }
    for (; i < n; i += stride) acc0 += s[i].x + s[i].y;
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc0 + acc1;
}

// Two more read patterns for the same ceiling: 64 contiguous bytes per lane
// per iteration (four 16-byte loads from one lane's own run), and 16-byte
// non-temporal loads (no reuse expected in the caches).
__global__ __launch_bounds__(256) void stream_read_wide_kernel(const double2* __restrict__ s,
                                                               double* __restrict__ out, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc0 = 0.0, acc1 = 0.0;
    for (; 4 * i + 3 < n; i += stride) {
        const double2 a = s[4 * i], b = s[4 * i + 1], c = s[4 * i + 2], e = s[4 * i + 3];
        acc0 += (a.x + b.x) + (c.x + e.x);
        acc1 += (a.y + b.y) + (c.y + e.y);
    }
    for (size_t j = 4 * i; j < n && j < 4 * i + 4; ++j) acc0 += s[j].x + s[j].y;
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc0 + acc1;
}

typedef double v2d __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void stream_read_nt_kernel(const double2* __restrict__ s2, double* __restrict__ out,
                                                             size_t n) {
    const v2d* s = reinterpret_cast<const v2d*>(s2);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc0 = 0.0, acc1 = 0.0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const v2d a = __builtin_nontemporal_load(s + i), b = __builtin_nontemporal_load(s + i + stride);
        const v2d c = __builtin_nontemporal_load(s + i + 2 * stride);
        const v2d e = __builtin_nontemporal_load(s + i + 3 * stride);
        acc0 += (a.x + b.x) + (c.x + e.x);
        acc1 += (a.y + b.y) + (c.y + e.y);
    }
    for (; i < n; i += stride) acc0 += s[i].x + s[i].y;
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc0 + acc1;
}

// Device math probes for the parity tests (sqrt, division, pow, pow6).
__global__ void probe_kernel(const double* x, const double* y, double* out, int n, int which) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (which) {
    case 0: out[i] = sqrt(x[i]); break;
    case 1: out[i] = x[i] / y[i]; break;
    case 2: out[i] = ddpow::pow_dd(x[i], y[i]); break;
    case 4: out[i] = pow(x[i], y[i]); break;
    case 3: out[i] = pow6_cr(x[i]); break;
    default: out[i] = 0.0;
    }
}

}  // namespace

hipError_t teno_faces(const Geom& g, const Teno& t, const double* U, double* fs, uint8_t* dec,
                      const Physics& ph, int occ, int gate_reuse, hipStream_t stream) {
    const int Kp = t.K - 1, Mp = t.M - 1;
    const uint32_t FQ = g.S * g.Q;
    if (g.S != 3 || (FQ != 3 && FQ != 6) || Mp < 1) return hipErrorInvalidValue;
    if (occ != 0 && occ != 4 && occ != 5 && occ != 6) return hipErrorInvalidValue;
    const uint32_t nblk = (g.N + 63) / 64, per_xcd = (nblk + 7) / 8;
    const dim3 grid(8 * per_xcd), block(256);
    const bool small = t.loc8 != nullptr;
    if (!small && t.loc16 == nullptr) return hipErrorInvalidValue;
    const size_t net_lds = (t.gate && !gate_reuse) ? ind::kNetInferBytes : 0;
    // small: 8-bit tile positions (tiles of <= 256 cells); O: occupancy target
#define MLR_TENO_S(KP, MP, FQV, G, O)                                                                       \
    hipLaunchKernelGGL((teno_kernel<KP, MP, FQV, uint8_t, 256, G, O>), grid, block, 256 * 32 + net_lds, stream, \
                       g, t, U, fs, dec, ph, per_xcd)
#define MLR_TENO_L(KP, MP, FQV, G)                                                                         \
    hipLaunchKernelGGL((teno_kernel<KP, MP, FQV, uint16_t, 1024, G, 0>), grid, block, 1024 * 32 + net_lds,   \
                       stream, g, t, U, fs, dec, ph, per_xcd)
#define MLR_TENO_G(KP, MP, FQV, G)                  \
    do {                                            \
        if (small) MLR_TENO_S(KP, MP, FQV, G, 0);   \
        else MLR_TENO_L(KP, MP, FQV, G);            \
    } while (0)
    // the gated kernels exist where the central stencil has the network's 12 members or more
#define MLR_TENO(KP, MP, FQV)                              \
    do {                                                   \
        if (!t.gate) MLR_TENO_G(KP, MP, FQV, 0);           \
        else if (gate_reuse) MLR_TENO_G(KP, MP, FQV, 2);   \
        else MLR_TENO_G(KP, MP, FQV, 1);                   \
    } while (0)
    if (t.gate && Mp < ind::kMembers - 1) return hipErrorInvalidValue;
    if (t.derived) {
        // derived face basis: plain TENO at p = 2 with the default stencil
        // size and 8-bit tile positions
        if (!(FQ == 3 && Kp == 5 && Mp == 11 && small) || t.cvert == nullptr || t.compact || t.gate)
            return hipErrorInvalidValue;
        const size_t lds = 256 * 32;
        if (occ == 4)
            hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 0, 4, false, true>), grid, block, lds, stream, g,
                               t, U, fs, dec, ph, per_xcd);
        else if (occ == 5)
            hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 0, 5, false, true>), grid, block, lds, stream, g,
                               t, U, fs, dec, ph, per_xcd);
        else if (occ == 0)
            hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 0, 0, false, true>), grid, block, lds, stream, g,
                               t, U, fs, dec, ph, per_xcd);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (t.compact) {
        // compact tables: p = 2 with the default stencil size, 8-bit tile
        // positions, the compiler's occupancy
        if (!(FQ == 3 && Kp == 5 && Mp == 11 && small) || t.fbc == nullptr) return hipErrorInvalidValue;
        const size_t lds = 256 * 32;
        if (t.gate && t.blk_mixed) {
            // split gated pass: the gated cells (linear operators, no tables),
            // then the ungated cells of the marked blocks on compact tables
            if (gate_reuse) MLR_TENO_S(5, 11, 3, 4, 0);
            else MLR_TENO_S(5, 11, 3, 3, 0);
            if (const hipError_t e = hipGetLastError(); e != hipSuccess) return e;
            hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 5, 0, true>), grid, block, lds, stream, g, t, U,
                               fs, dec, ph, per_xcd);
        } else if (!t.gate) {
            if (occ == 5)
                hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 0, 5, true>), grid, block, lds, stream, g, t,
                                   U, fs, dec, ph, per_xcd);
            else if (occ == 6)
                hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 0, 6, true>), grid, block, lds, stream, g, t,
                                   U, fs, dec, ph, per_xcd);
            else
                hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 0, 0, true>), grid, block, lds, stream, g, t,
                                   U, fs, dec, ph, per_xcd);
        } else if (gate_reuse) {
            hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 2, 0, true>), grid, block, lds, stream, g, t, U,
                               fs, dec, ph, per_xcd);
        } else {
            hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 1, 0, true>), grid, block, lds + net_lds, stream,
                               g, t, U, fs, dec, ph, per_xcd);
        }
        return hipGetLastError();
    }
    // the default stencil size (factor 2: M = 2K) is compiled with the member
    // loop unrolled, and at p = 2 in three occupancy targets; other sizes
    // take the runtime loop
    if (FQ == 3) {
        switch (Kp) {
        case 2: if (Mp == 5) MLR_TENO_G(2, 5, 3, 0); else MLR_TENO_G(2, 0, 3, 0); break;
        case 5:
            if (Mp == 11 && small && t.gate && t.blk_mixed) {
                // split gated pass: the gated cells (modes 3 / 4, no
                // full-TENO code), then the ungated ones of the marked
                // blocks (mode 5). The evaluate pass targets 6 waves per
                // SIMD (77 VGPRs, no spill; the compiler's own choice, 75
                // VGPRs + 8 AGPRs, gives 5); the others run at the
                // compiler's occupancy (the tune's target does not apply)
                if (gate_reuse) MLR_TENO_S(5, 11, 3, 4, 0);
                else MLR_TENO_S(5, 11, 3, 3, 6);
                if (const hipError_t e = hipGetLastError(); e != hipSuccess) return e;
                hipLaunchKernelGGL((teno_kernel<5, 11, 3, uint8_t, 256, 5, 0>), grid, block, 256 * 32, stream, g, t, U,
                                   fs, dec, ph, per_xcd);
            } else if (Mp == 11 && small) {
                const int gm = !t.gate ? 0 : (gate_reuse ? 2 : 1);
#define MLR_TENO_O(GMV)                                       \
    do {                                                      \
        if (occ == 5) MLR_TENO_S(5, 11, 3, GMV, 5);           \
        else if (occ == 6) MLR_TENO_S(5, 11, 3, GMV, 6);      \
        else MLR_TENO_S(5, 11, 3, GMV, 0);                    \
    } while (0)
                if (gm == 0) MLR_TENO_O(0);
                else if (gm == 1) MLR_TENO_O(1);
                else MLR_TENO_O(2);
#undef MLR_TENO_O
            } else if (Mp == 11) {
                MLR_TENO(5, 11, 3);
            } else {
                MLR_TENO(5, 0, 3);
            }
            break;
        case 9: MLR_TENO(9, 0, 3); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (Kp) {
        case 2: MLR_TENO_G(2, 0, 6, 0); break;
        case 5: MLR_TENO(5, 0, 6); break;
        case 9:
            if (Mp == 19 && !t.gate && (occ == 4 || occ == 5)) {  // plain p = 3 at an occupancy target
                if (small && occ == 4) MLR_TENO_S(9, 19, 6, 0, 4);
                else if (small) MLR_TENO_S(9, 19, 6, 0, 5);
                else if (occ == 4)
                    hipLaunchKernelGGL((teno_kernel<9, 19, 6, uint16_t, 1024, 0, 4>), grid, block, 1024 * 32, stream,
                                       g, t, U, fs, dec, ph, per_xcd);
                else
                    hipLaunchKernelGGL((teno_kernel<9, 19, 6, uint16_t, 1024, 0, 5>), grid, block, 1024 * 32, stream,
                                       g, t, U, fs, dec, ph, per_xcd);
            } else if (Mp == 19) {
                MLR_TENO(9, 19, 6);
            } else {
                MLR_TENO(9, 0, 6);
            }
            break;
        default: return hipErrorInvalidValue;
        }
    }
#undef MLR_TENO_S
#undef MLR_TENO_L
#undef MLR_TENO_G
#undef MLR_TENO
    return hipGetLastError();
}

hipError_t flux_update(const Geom& g, const StageArgs& a, const Physics& ph, uint32_t bs, hipStream_t stream) {
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    const uint32_t nblk = (g.N + bs - 1) / bs, per_xcd = (nblk + 7) / 8;
    const dim3 grid(8 * per_xcd), block(bs);
    const bool pp = a.pp && a.op != kRhsOnly;
#define MLR_FU(RSV)                                                                                           \
    do {                                                                                                      \
        if (pp) hipLaunchKernelGGL((flux_update_kernel<RSV, true>), grid, block, 0, stream, g, a, ph, per_xcd); \
        else hipLaunchKernelGGL((flux_update_kernel<RSV, false>), grid, block, 0, stream, g, a, ph, per_xcd);   \
    } while (0)
    switch (a.riemann) {
    case 0: MLR_FU(0); break;
    case 1: MLR_FU(1); break;
    case 2: MLR_FU(2); break;
    case 3: MLR_FU(3); break;
    default: return hipErrorInvalidValue;
    }
#undef MLR_FU
    return hipGetLastError();
}

hipError_t rhs_parts(const Geom& g, const StageArgs& a, const Physics& ph, int parts, uint32_t bs,
                     hipStream_t stream) {
    if (parts == 3) return flux_update(g, a, ph, bs, stream);
    if ((parts != 1 && parts != 2) || a.op != kRhsOnly) return hipErrorInvalidValue;
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    const uint32_t nblk = (g.N + bs - 1) / bs, per_xcd = (nblk + 7) / 8;
    const dim3 grid(8 * per_xcd), block(bs);
#define MLR_PARTS(RSV)                                                                                        \
    do {                                                                                                      \
        if (parts == 1) hipLaunchKernelGGL((rhs_parts_kernel<RSV, 1>), grid, block, 0, stream, g, a, ph, per_xcd); \
        else hipLaunchKernelGGL((rhs_parts_kernel<RSV, 2>), grid, block, 0, stream, g, a, ph, per_xcd);          \
    } while (0)
    switch (a.riemann) {
    case 0: MLR_PARTS(0); break;
    case 1: MLR_PARTS(1); break;
    case 2: MLR_PARTS(2); break;
    case 3: MLR_PARTS(3); break;
    default: return hipErrorInvalidValue;
    }
#undef MLR_PARTS
    return hipGetLastError();
}

hipError_t spectral_radius(const Geom& g, const double* Un, const double* prims, const Physics& ph,
                           double* lam_cell, unsigned long long* lam_blk, uint32_t bs, hipStream_t stream) {
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    const uint32_t nblk = (g.N + bs - 1) / bs, per_xcd = (nblk + 7) / 8;
    hipLaunchKernelGGL(spectral_kernel, dim3(8 * per_xcd), dim3(bs), 0, stream, g, Un, prims, ph, lam_cell, lam_blk,
                       per_xcd);
    return hipGetLastError();
}

hipError_t finalize_dt(const unsigned long long* lam_blk, uint32_t n_blk, double cfl, double fixed_dt, double* dt,
                       double* dt_hist, int64_t hist_cap, int64_t* step, hipStream_t stream) {
    hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(1024), 0, stream, lam_blk, n_blk, cfl, fixed_dt, dt, dt_hist,
                       hist_cap, step);
    return hipGetLastError();
}

hipError_t teno_batch(const CaseBufs* cases, int n_cases, uint32_t total_blocks, int in_role, int K, int M, int FQ,
                      bool small, bool gated, int occ, int gate_reuse, bool split, bool derived, const Physics& ph,
                      hipStream_t stream) {
    const int Kp = K - 1, Mp = M - 1;
    const size_t net_lds = (gated && !gate_reuse) ? ind::kNetInferBytes : 0;
    if (n_cases < 1 || (FQ != 3 && FQ != 6) || Mp < 1) return hipErrorInvalidValue;
    const uint32_t per_xcd = (total_blocks + 7) / 8;
    const dim3 grid(8 * per_xcd), block(256);
    if (derived) {
        // derived face basis: plain p = 2 with the default stencil size and
        // 8-bit tile positions in every case (each case's Teno::cvert set)
        if (!(FQ == 3 && Kp == 5 && Mp == 11 && small) || gated) return hipErrorInvalidValue;
        if (occ == 4)
            hipLaunchKernelGGL((teno_batch_kernel<5, 11, 3, uint8_t, 256, 0, 4, true>), grid, block, 256 * 32,
                               stream, cases, n_cases, in_role, total_blocks, ph, per_xcd);
        else if (occ == 5)
            hipLaunchKernelGGL((teno_batch_kernel<5, 11, 3, uint8_t, 256, 0, 5, true>), grid, block, 256 * 32,
                               stream, cases, n_cases, in_role, total_blocks, ph, per_xcd);
        else if (occ == 0)
            hipLaunchKernelGGL((teno_batch_kernel<5, 11, 3, uint8_t, 256, 0, 0, true>), grid, block, 256 * 32,
                               stream, cases, n_cases, in_role, total_blocks, ph, per_xcd);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (occ != 0 && occ != 5 && occ != 6) return hipErrorInvalidValue;
#define MLR_TB_S(KP, MP, FQV, G, O)                                                                             \
    hipLaunchKernelGGL((teno_batch_kernel<KP, MP, FQV, uint8_t, 256, G, O>), grid, block, 256 * 32 + net_lds,   \
                       stream, cases, n_cases, in_role, total_blocks, ph, per_xcd)
#define MLR_TB_G(KP, MP, FQV, G)                                                                               \
    do {                                                                                                       \
        if (small)                                                                                             \
            MLR_TB_S(KP, MP, FQV, G, 0);                                                                       \
        else                                                                                                   \
            hipLaunchKernelGGL((teno_batch_kernel<KP, MP, FQV, uint16_t, 1024, G, 0>), grid, block,              \
                               1024 * 32 + net_lds, stream, cases, n_cases, in_role, total_blocks, ph, per_xcd); \
    } while (0)
#define MLR_TB(KP, MP, FQV)                              \
    do {                                                 \
        if (!gated) MLR_TB_G(KP, MP, FQV, 0);            \
        else if (gate_reuse) MLR_TB_G(KP, MP, FQV, 2);   \
        else MLR_TB_G(KP, MP, FQV, 1);                   \
    } while (0)
    if (gated && Mp < ind::kMembers - 1) return hipErrorInvalidValue;
    if (FQ == 3 && Kp == 5 && Mp == 11 && small && gated && split) {
        // split gated pass (teno_faces' modes 3 / 4, then 5)
        if (gate_reuse) MLR_TB_S(5, 11, 3, 4, 0);
        else MLR_TB_S(5, 11, 3, 3, 6);  // the evaluate pass at 6 waves per SIMD, as teno_faces
        if (const hipError_t e = hipGetLastError(); e != hipSuccess) return e;
        hipLaunchKernelGGL((teno_batch_kernel<5, 11, 3, uint8_t, 256, 5, 0>), grid, block, 256 * 32, stream, cases,
                           n_cases, in_role, total_blocks, ph, per_xcd);
    } else if (FQ == 3 && Kp == 5 && Mp == 11 && small) {
        const int gm = !gated ? 0 : (gate_reuse ? 2 : 1);
#define MLR_TB_O(GMV)                                     \
    do {                                                  \
        if (occ == 5) MLR_TB_S(5, 11, 3, GMV, 5);         \
        else if (occ == 6) MLR_TB_S(5, 11, 3, GMV, 6);    \
        else MLR_TB_S(5, 11, 3, GMV, 0);                  \
    } while (0)
        if (gm == 0) MLR_TB_O(0);
        else if (gm == 1) MLR_TB_O(1);
        else MLR_TB_O(2);
#undef MLR_TB_O
    } else if (FQ == 3 && Kp == 5 && Mp == 11) MLR_TB(5, 11, 3);
    else if (FQ == 3 && Kp == 5) MLR_TB(5, 0, 3);
    else if (FQ == 3 && Kp == 2) MLR_TB_G(2, 0, 3, 0);
    else if (FQ == 6 && Kp == 9) MLR_TB(9, 0, 6);
    else return hipErrorInvalidValue;
#undef MLR_TB_S
#undef MLR_TB_G
#undef MLR_TB
    return hipGetLastError();
}

hipError_t flux_batch(const CaseBufs* cases, int n_cases, uint32_t total_blocks, const BatchStage& st, int riemann,
                      const Physics& ph, uint32_t bs, hipStream_t stream) {
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    const uint32_t per_xcd = (total_blocks + 7) / 8;
    const dim3 grid(8 * per_xcd), block(bs);
#define MLR_FB(RSV)                                                                                   \
    do {                                                                                              \
        if (st.pp)                                                                                    \
            hipLaunchKernelGGL((flux_batch_kernel<RSV, true>), grid, block, 0, stream, cases, n_cases, st, \
                               total_blocks, ph, per_xcd);                                            \
        else                                                                                          \
            hipLaunchKernelGGL((flux_batch_kernel<RSV, false>), grid, block, 0, stream, cases, n_cases, st, \
                               total_blocks, ph, per_xcd);                                            \
    } while (0)
    switch (riemann) {
    case 0: MLR_FB(0); break;
    case 1: MLR_FB(1); break;
    case 2: MLR_FB(2); break;
    case 3: MLR_FB(3); break;
    default: return hipErrorInvalidValue;
    }
#undef MLR_FB
    return hipGetLastError();
}

hipError_t spectral_batch(const CaseBufs* cases, int n_cases, uint32_t total_blocks, const Physics& ph, uint32_t bs,
                          hipStream_t stream) {
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    const uint32_t per_xcd = (total_blocks + 7) / 8;
    hipLaunchKernelGGL(spectral_batch_kernel, dim3(8 * per_xcd), dim3(bs), 0, stream, cases, n_cases, total_blocks,
                       ph, per_xcd);
    return hipGetLastError();
}

hipError_t finalize_batch(const CaseBufs* cases, int n_cases, double cfl, double fixed_dt, int64_t hist_cap,
                          uint32_t bs, hipStream_t stream) {
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(finalize_batch_kernel, dim3(n_cases), dim3(256), 0, stream, cases, n_cases, cfl, fixed_dt,
                       hist_cap, bs);
    return hipGetLastError();
}

namespace {
template <int RS, bool TENO>
hipError_t launch_step_persistent(const CaseBufs* cases, int n_cases, uint32_t teno_total, uint32_t flux_total,
                                  const PersistentStep& ps, const Physics& ph, hipStream_t stream, int* grid_out) {
    const void* fn = reinterpret_cast<const void*>(&ensemble_step_kernel<RS, TENO>);
    const size_t lds = TENO ? 256 * 32 : 0;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds); e != hipSuccess) return e;
    if (hipError_t e = hipGetDevice(&dev); e != hipSuccess) return e;
    if (hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev); e != hipSuccess)
        return e;
    const int grid = per_cu * cus;  // every block resident at once: the grid barrier needs it
    if (grid < 8) return hipErrorInvalidConfiguration;  // each XCD's range needs a block
    if (grid_out) *grid_out = grid;
    const CaseBufs* c = cases;
    int n = n_cases;
    uint32_t tt = teno_total, ft = flux_total;
    PersistentStep p = ps;
    Physics h = ph;
    void* args[] = {&c, &n, &tt, &ft, &p, &h};
    return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), args, (unsigned)lds, stream);
}
}  // namespace

hipError_t ensemble_step(const CaseBufs* cases, int n_cases, uint32_t teno_total, uint32_t flux_total,
                         const PersistentStep& ps, int riemann, bool teno, const Physics& ph, hipStream_t stream,
                         int* grid_out) {
    if (n_cases < 1 || ps.n_stages < 1 || ps.n_stages > 4) return hipErrorInvalidValue;
#define MLR_STEP(RSV)                                                                                          \
    return teno ? launch_step_persistent<RSV, true>(cases, n_cases, teno_total, flux_total, ps, ph, stream, grid_out) \
                : launch_step_persistent<RSV, false>(cases, n_cases, teno_total, flux_total, ps, ph, stream, grid_out)
    switch (riemann) {
    case 0: MLR_STEP(0);
    case 1: MLR_STEP(1);
    case 2: MLR_STEP(2);
    case 3: MLR_STEP(3);
    default: return hipErrorInvalidValue;
    }
#undef MLR_STEP
}

hipError_t viscous_grad(const Geom& g, const double* U, const Physics& ph, double* vp, double* vg, uint32_t bs,
                        hipStream_t stream) {
    if (bs != 64 && bs != 128 && bs != 256) return hipErrorInvalidValue;
    const uint32_t nblk = (g.N + bs - 1) / bs, per_xcd = (nblk + 7) / 8;
    hipLaunchKernelGGL(viscous_grad_kernel, dim3(8 * per_xcd), dim3(bs), 0, stream, g, U, ph, vp, vg, per_xcd);
    return hipGetLastError();
}

hipError_t copy_state(const double* src, double* dst, uint32_t N, hipStream_t stream) {
    const size_t n = 4 * (size_t)N;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

uint32_t stream_probe_threads(size_t bytes) {
    const size_t n = bytes / 16;
    const size_t blocks = (n + 255) / 256;
    return (uint32_t)(blocks < 16384 ? blocks : 16384) * 256u;
}

hipError_t stream_probe(int kind, const void* src, void* dst, size_t bytes, hipStream_t stream) {
    const size_t n = bytes / 16;
    const uint32_t threads = stream_probe_threads(bytes);
    if (threads == 0) return hipErrorInvalidValue;
    const dim3 grid(threads / 256), block(256);
    switch (kind) {
    case 0:
        hipLaunchKernelGGL(stream_copy_kernel, grid, block, 0, stream, static_cast<const double2*>(src),
                           static_cast<double2*>(dst), n);
        break;
    case 1:
        hipLaunchKernelGGL(stream_read_kernel, grid, block, 0, stream, static_cast<const double2*>(src),
                           static_cast<double*>(dst), n);
        break;
    case 2: return hipMemcpyAsync(dst, src, bytes / 16 * 16, hipMemcpyDeviceToDevice, stream);
    case 3:
        hipLaunchKernelGGL(stream_read_wide_kernel, grid, block, 0, stream, static_cast<const double2*>(src),
                           static_cast<double*>(dst), n);
        break;
    case 4:
        hipLaunchKernelGGL(stream_read_nt_kernel, grid, block, 0, stream, static_cast<const double2*>(src),
                           static_cast<double*>(dst), n);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t probe_math(const double* x, const double* y, double* out, int n, int which, hipStream_t stream) {
    hipLaunchKernelGGL(probe_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, x, y, out, n, which);
    return hipGetLastError();
}

}  // namespace mlr::dev
