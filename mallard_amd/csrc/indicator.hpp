// Learned troubled-cell indicator: per-cell stencil features and a
// 48-64-64-1 MLP (GELU, GELU, sigmoid) with BF16 weights and activations and
// FP32 accumulation on the CDNA matrix cores (v_mfma_f32_16x16x16_bf16).
//
// Features of cell c, from the 12 members of its central stencil (target
// first). For members m = 1..11 the raw differences, in FP64,
//   d = (rho_m / rho_c - 1, (u_m - u_c) / a_c, (v_m - v_c) / a_c, p_m / p_c - 1),
// a_c = sqrt(gamma p_c / rho_c), are divided per variable by their largest
// magnitude s_v over the stencil (0 when s_v = 0), so the pattern is scale
// free; member 0's four slots carry the scale, ilogb(s_v) / 32 clipped below
// at -2 (-2.5 when s_v = 0). Values are rounded to FP32 then BF16 (nearest
// even). A cell with a non-physical member (or itself) has no features and
// is never judged smooth.
//
// MFMA mapping (one wave = 16 cells, MFMA column n = lane % 16, row group
// g = lane / 16): every layer is computed transposed, D = W^T X^T, so the
// accumulator a lane receives (rows 4g..4g+3 of its column) is already the
// B operand layout of the next layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace mlr::dev::ind {

constexpr int kFeat = 48, kHid = 64, kMembers = 12;
// master parameter vector (FP32): W1[48][64], b1[64], W2[64][64], b2[64], w3[64], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeat * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
static_assert(kOffB3 + 1 == kNParams, "parameter layout");
// packed MFMA A operands: W1p[nt 4][kc 3][lane 64], W2p[nt 4][kc 4][lane 64], 4 bf16 each
constexpr int kW1pVec = 4 * 3 * 64, kW2pVec = 4 * 4 * 64;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f32_to_bf16(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7FFFu + ((u >> 16) & 1u);                                             // nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// tanh form of GELU and its derivative, FP32
__device__ inline float gelu(float x) {
    const float k = 0.7978845608028654f;
    const float t = tanhf(k * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.0f + t);
}
__device__ inline float gelu_d(float x) {
    const float k = 0.7978845608028654f;
    const float t = tanhf(k * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k * (1.0f + 3.0f * 0.044715f * x * x);
}

struct CellRef {
    double rho, u, v, p, a;
    bool ok;
};

__device__ inline double pressure_of(double gamma, const double* U) {
    return (gamma - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
}

__device__ inline CellRef cell_ref(double gamma, const double* U) {
    CellRef r;
    r.rho = U[0];
    r.u = U[1] / U[0];
    r.v = U[2] / U[0];
    r.p = pressure_of(gamma, U);
    r.ok = r.rho > 0.0 && r.p > 0.0;
    r.a = r.ok ? sqrt(gamma * r.p / r.rho) : 1.0;
    return r;
}

__device__ inline uint16_t feat_bf16(double x) { return f32_to_bf16((float)x); }

// raw differences of member state Um against the cell; false when Um is not physical
__device__ inline bool member_diff(double gamma, const CellRef& r, const double* Um, double* d) {
    const double rho = Um[0];
    const double u = Um[1] / rho, v = Um[2] / rho;
    const double p = pressure_of(gamma, Um);
    d[0] = rho / r.rho - 1.0;
    d[1] = (u - r.u) / r.a;
    d[2] = (v - r.v) / r.a;
    d[3] = p / r.p - 1.0;
    return rho > 0.0 && p > 0.0 && d[0] == d[0] && d[1] == d[1] && d[2] == d[2] && d[3] == d[3];
}

// scale slot of member 0
__device__ inline double scale_feature(double s) { return s > 0.0 ? fmax(-2.0, (double)ilogb(s) / 32.0) : -2.5; }

// pattern slot of members 1..11
__device__ inline double pattern_feature(double d, double s) { return s > 0.0 ? d / s : 0.0; }

__device__ inline v4s pack4(const uint16_t* f) {
    v4s x;
    x[0] = (short)f[0];
    x[1] = (short)f[1];
    x[2] = (short)f[2];
    x[3] = (short)f[3];
    return x;
}

// Forward pass of one wave. xb[kc] holds, for cell lane % 16, the features
// kc*16 + 4g .. kc*16 + 4g + 3 (that is member 4 kc + g). Every lane returns
// the logit of cell lane % 16. When pre1 / h1 / pre2 are given the lane also
// stores its 16 pre-activations and hidden values (hidden index
// nt*16 + 4g + r at [nt*4 + r]) for the backward pass.
__device__ inline float mlp_wave(const Net& net, const v4s* xb, int lane, float* pre1 = nullptr,
                                 float* h1s = nullptr, float* pre2 = nullptr) {
    const int g = lane >> 4;
    v4s h1[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < 3; ++kc)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(static_cast<const v4s*>(net.W1p)[(nt * 3 + kc) * 64 + lane],
                                                           xb[kc], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float pre = acc[r] + net.b1[nt * 16 + 4 * g + r];
            const uint16_t hb = f32_to_bf16(gelu(pre));
            h1[nt][r] = (short)hb;
            if (pre1) {
                pre1[nt * 4 + r] = pre;
                h1s[nt * 4 + r] = bf16_to_f32(hb);
            }
        }
    }
    float z = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < 4; ++kc)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(static_cast<const v4s*>(net.W2p)[(nt * 4 + kc) * 64 + lane],
                                                           h1[kc], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float pre = acc[r] + net.b2[nt * 16 + 4 * g + r];
            if (pre2) pre2[nt * 4 + r] = pre;
            z += net.w3[nt * 16 + 4 * g + r] * gelu(pre);
        }
    }
    // the 4 lanes of a column hold the 4 row groups of the same cell
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    return z + *net.b3;
}

}  // namespace mlr::dev::ind
