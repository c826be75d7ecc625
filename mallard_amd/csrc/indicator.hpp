// Learned troubled-cell indicator: per-cell stencil features and a
// 48-64-64-1 MLP (GELU, GELU, sigmoid) with BF16 weights and activations and
// FP32 accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16,
// 32 products per lane group and output).
//
// Features of cell c, from the 12 members of its central stencil (target
// first). For members m = 1..11 the raw differences, in FP64,
//   d = (rho_m ir_c - 1, (u_m - u_c) ia_c, (v_m - v_c) ia_c, p_m ip_c - 1),
// with the cell's reciprocals ir_c = 1 / rho_c, ia_c = 1 / a_c,
// ip_c = 1 / p_c, a_c = sqrt(gamma p_c ir_c), and the member's velocity and
// pressure formed with its own 1 / rho_m, are multiplied per variable by the
// reciprocal of their largest magnitude s_v over the stencil (0 when s_v = 0),
// so the pattern is scale free; member 0's four slots carry the scale,
// ilogb(s_v) / 32 clipped below at -2 (-2.5 when s_v = 0). Values are rounded
// to FP32 then BF16 (nearest even). A cell with a non-physical member (or
// itself) has no features and is never judged smooth.
//
// MFMA mapping (one wave = 16 cells, MFMA column n = lane % 16, row group
// g = lane / 16): every layer is computed transposed, D = W^T X^T, so the
// accumulator a lane receives (rows 4g..4g+3 of its column) is already half
// of a B operand of the next layer. A 16x16x32 operand takes 8 products per
// lane: the two 4-element groups a lane already holds (features of members
// g and 4 + g, then of 8 + g and zeros; hidden units 32 kk + 4g + r and
// 32 kk + 16 + 4g + r) side by side, and the weights are packed in that
// order. Both operands pair element e of lane group g with the same product
// index, so the order within the 32 only permutes the sum.
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
// packed MFMA A operands: W1p[nt 4][kk 2][lane 64], W2p[nt 4][kk 2][lane 64], 8 bf16 each
constexpr int kW1pVec = 4 * 2 * 64, kW2pVec = 4 * 2 * 64;
constexpr int kOperandBytes = 16;
// the inference part of the packed network (W1p, W2p, b1, b2, w3, b3 from
// the buffer's start), padded to 16 bytes: what a block stages in LDS
constexpr int kNetInferBytes = ((kW1pVec + kW2pVec) * kOperandBytes + (3 * kHid + 1) * 4 + 15) / 16 * 16;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));

__host__ __device__ inline uint16_t f32_to_bf16(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7FFFu + ((u >> 16) & 1u);                                             // nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Two FP32 values to a bf16 pair, round to nearest even, lo in the low half
// (the hardware converter; the same rounding as f32_to_bf16 for every
// non-NaN input).
__device__ inline uint32_t bf16x2(float lo, float hi) {
    const v2f f = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, v2bf));
}
__device__ inline float bf16_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ inline float bf16_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xFFFF0000u); }

// FP64 -> FP32 as a step of its own: features are rounded to FP32, then to
// bf16 (kept opaque so the two roundings are not merged into one).
__device__ inline float to_f32(double x) {
    float r;
    asm("v_cvt_f32_f64 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// four features (FP64) as the bf16 MFMA operand of one lane
__device__ inline v4s pack_feats(const double* x) {
    const v2u w = {bf16x2(to_f32(x[0]), to_f32(x[1])), bf16x2(to_f32(x[2]), to_f32(x[3]))};
    return __builtin_bit_cast(v4s, w);
}

// tanh form of GELU, 0.5 x (1 + tanh(u)) = x sigmoid(2u) with
// u = k (x + 0.044715 x^3), k = sqrt(2 / pi), in FP32 with the hardware exp2
// and reciprocal (about 1 ulp each): sigmoid(2u) = 1 / (1 + 2^t),
// t = -2u log2(e) = x (-2k log2(e) - 2k 0.044715 log2(e) x^2). Its derivative
// is sigmoid(2u) + 2 x sigmoid(2u) (1 - sigmoid(2u)) u'.
__device__ inline float gelu_sig(float x) {
    const float t = x * __builtin_fmaf(-0.1029432395800235f, x * x, -2.302208198144325f);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
__device__ inline float gelu(float x) { return x * gelu_sig(x); }
__device__ inline float gelu_d(float x) {
    const float s = gelu_sig(x);
    const float du = 0.7978845608028654f * __builtin_fmaf(0.134145f, x * x, 1.0f);
    return s + 2.0f * x * s * (1.0f - s) * du;
}

struct CellRef {
    double u, v, ir, ia, ip;
    bool ok;
};

__device__ inline CellRef cell_ref(double gamma, const double* U) {
    CellRef r;
    r.ir = 1.0 / U[0];
    r.u = U[1] * r.ir;
    r.v = U[2] * r.ir;
    const double p = (gamma - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) * r.ir);
    r.ok = U[0] > 0.0 && p > 0.0;
    r.ia = r.ok ? 1.0 / sqrt(gamma * p * r.ir) : 1.0;
    r.ip = 1.0 / p;
    return r;
}

// cell_ref evaluated by the 4 lanes of the cell's MFMA column (lanes col,
// col + 16, col + 32, col + 48; gq = lane / 16): its two reciprocals 1 / a_c
// and 1 / p_c are taken by lanes gq = 0 and 1 and shared, so each lane runs
// one division fewer. The operations and operands are cell_ref's, so are the
// results.
__device__ inline CellRef cell_ref_col(double gamma, const double* U, int gq, int col) {
    CellRef r;
    r.ir = 1.0 / U[0];
    r.u = U[1] * r.ir;
    r.v = U[2] * r.ir;
    const double p = (gamma - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) * r.ir);
    r.ok = U[0] > 0.0 && p > 0.0;
    const double rec = 1.0 / (gq == 0 ? sqrt(gamma * p * r.ir) : p);
    const double ra = __shfl(rec, col), rp = __shfl(rec, col + 16);
    r.ia = r.ok ? ra : 1.0;
    r.ip = rp;
    return r;
}

// raw differences of member state Um against the cell; false when Um is not physical
__device__ inline bool member_diff(double gamma, const CellRef& r, const double* Um, double* d) {
    const double rho = Um[0];
    const double irm = 1.0 / rho;
    const double u = Um[1] * irm, v = Um[2] * irm;
    const double p = (gamma - 1.0) * (Um[3] - 0.5 * (Um[1] * Um[1] + Um[2] * Um[2]) * irm);
    d[0] = rho * r.ir - 1.0;
    d[1] = (u - r.u) * r.ia;
    d[2] = (v - r.v) * r.ia;
    d[3] = p * r.ip - 1.0;
    return rho > 0.0 && p > 0.0 && isfinite(d[0]) && isfinite(d[1]) && isfinite(d[2]) && isfinite(d[3]);
}

// scale slot of member 0
__device__ inline double scale_feature(double s) { return s > 0.0 ? fmax(-2.0, (double)ilogb(s) / 32.0) : -2.5; }

// reciprocal of the stencil-wide maximum s (0 when s = 0) and the pattern
// slot of members 1..11
__device__ inline double inv_scale(double s) { return s > 0.0 ? 1.0 / s : 0.0; }
__device__ inline double pattern_feature(double d, double is) { return d * is; }

__device__ inline v4s pack4(const uint16_t* f) {
    v4s x;
    x[0] = (short)f[0];
    x[1] = (short)f[1];
    x[2] = (short)f[2];
    x[3] = (short)f[3];
    return x;
}

// Block-wide copy of the network's inference part into LDS (the caller
// synchronises before use) and the view of that copy.
__device__ inline void stage_net(const Net& net, unsigned char* lds) {
    const uint4* src = static_cast<const uint4*>(net.W1p);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int i = threadIdx.x; i < kNetInferBytes / 16; i += blockDim.x) dst[i] = src[i];
}
__device__ inline Net lds_net(const Net& net, const unsigned char* lds) {
    Net n = net;
    n.W1p = lds;
    n.W2p = lds + kW1pVec * kOperandBytes;
    n.b1 = reinterpret_cast<const float*>(lds + (kW1pVec + kW2pVec) * kOperandBytes);
    n.b2 = n.b1 + kHid;
    n.w3 = n.b2 + kHid;
    n.b3 = n.w3 + kHid;
    return n;
}

// the 16x16x32 BF16 product with its two 4-element halves given apart
__device__ inline v4f mfma32(v8s a, v4s b_lo, v4s b_hi, v4f acc) {
    const v8s b = {b_lo[0], b_lo[1], b_lo[2], b_lo[3], b_hi[0], b_hi[1], b_hi[2], b_hi[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc,
                                                  0, 0, 0);
}

// Forward pass of one wave. xb[kc] holds, for cell lane % 16, the features
// kc*16 + 4g .. kc*16 + 4g + 3 (that is member 4 kc + g). Every lane returns
// the logit of cell lane % 16. The biases enter as the accumulators' initial
// values. When pre1 / h1 / pre2 are given the lane also stores its 16
// pre-activations and hidden values (hidden index nt*16 + 4g + r at
// [nt*4 + r]) for the backward pass.
__device__ inline float mlp_wave(const Net& net, const v4s* xb, int lane, float* pre1 = nullptr,
                                 float* h1s = nullptr, float* pre2 = nullptr) {
    const int g = lane >> 4;
    const float4* B1 = reinterpret_cast<const float4*>(net.b1);
    const float4* B2 = reinterpret_cast<const float4*>(net.b2);
    const float4* W3 = reinterpret_cast<const float4*>(net.w3);
    v4s h1[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float4 bb = B1[nt * 4 + g];
        v4f acc = {bb.x, bb.y, bb.z, bb.w};
        const v8s* W1 = static_cast<const v8s*>(net.W1p);
        acc = mfma32(W1[(nt * 2 + 0) * 64 + lane], xb[0], xb[1], acc);
        acc = mfma32(W1[(nt * 2 + 1) * 64 + lane], xb[2], v4s{0, 0, 0, 0}, acc);
        const v2u w = {bf16x2(gelu(acc[0]), gelu(acc[1])), bf16x2(gelu(acc[2]), gelu(acc[3]))};
        h1[nt] = __builtin_bit_cast(v4s, w);
        if (pre1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) pre1[nt * 4 + r] = acc[r];
            h1s[nt * 4 + 0] = bf16_lo(w[0]);
            h1s[nt * 4 + 1] = bf16_hi(w[0]);
            h1s[nt * 4 + 2] = bf16_lo(w[1]);
            h1s[nt * 4 + 3] = bf16_hi(w[1]);
        }
    }
    float z = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float4 bb = B2[nt * 4 + g];
        v4f acc = {bb.x, bb.y, bb.z, bb.w};
        const v8s* W2 = static_cast<const v8s*>(net.W2p);
        acc = mfma32(W2[(nt * 2 + 0) * 64 + lane], h1[0], h1[1], acc);
        acc = mfma32(W2[(nt * 2 + 1) * 64 + lane], h1[2], h1[3], acc);
        const float4 ww = W3[nt * 4 + g];
        if (pre2)
#pragma unroll
            for (int r = 0; r < 4; ++r) pre2[nt * 4 + r] = acc[r];
        z = __builtin_fmaf(ww.x, gelu(acc[0]), z);
        z = __builtin_fmaf(ww.y, gelu(acc[1]), z);
        z = __builtin_fmaf(ww.z, gelu(acc[2]), z);
        z = __builtin_fmaf(ww.w, gelu(acc[3]), z);
    }
    // the 4 lanes of a column hold the 4 row groups of the same cell
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    return z + *net.b3;
}

}  // namespace mlr::dev::ind
