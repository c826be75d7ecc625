// Learned troubled-cell indicator kernels (gfx950): feature extraction from a
// solver state with TENO labels, batched BF16 MFMA inference, and training by
// this file's own forward/backward kernels (per-block partial gradients in a
// fixed sample order, reduced in a fixed block order, Adam in FP32).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device.hpp"
#include "indicator.hpp"

namespace mlr::dev {

using namespace ind;

namespace {

// stencil member i (0-based after the target) of stencil 0 of cell c, as a
// cell id, read through the block tile tables
__device__ inline int32_t central_member(const Teno& t, uint32_t c, int i) {
    const int Mp = t.M - 1;
    const size_t o = teno_mem_off(c >> 6, Mp, i, c & 63u, 0);
    const uint32_t pos = t.loc8 ? (uint32_t)t.loc8[o] : (uint32_t)t.loc16[o];
    return t.blk_ids[t.blk_off[c >> 6] + pos];
}

// One thread per cell: the 48 features (bf16) and the TENO label (1 when the
// central stencil was selected for all 4 variables, -1 for a cell without
// features) at row base + c.
__global__ __launch_bounds__(256) void features_kernel(Geom g, Teno t, const double* __restrict__ U,
                                                       const uint8_t* __restrict__ dec, double gamma,
                                                       uint16_t* __restrict__ X, float* __restrict__ y,
                                                       uint64_t base) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const CellRef r = cell_ref(gamma, U + 4 * (size_t)c);
    double d[kMembers][4], smax[4] = {0.0, 0.0, 0.0, 0.0};
    bool ok = r.ok;
    for (int m = 1; m < kMembers && ok; ++m) {
        ok = member_diff(gamma, r, U + 4 * (size_t)central_member(t, c, m - 1), d[m]);
        for (int v = 0; v < 4; ++v) smax[v] = fmax(smax[v], fabs(d[m][v]));
    }
    uint16_t* row = X + (base + c) * kFeat;
    auto put = [&](int m, const double* x) {
        const v4s p = pack_feats(x);
        for (int v = 0; v < 4; ++v) row[4 * m + v] = (uint16_t)p[v];
    };
    double x[4], is[4];
    for (int v = 0; v < 4; ++v) {
        x[v] = ok ? scale_feature(smax[v]) : 0.0;
        is[v] = inv_scale(smax[v]);
    }
    put(0, x);
    for (int m = 1; m < kMembers; ++m) {
        for (int v = 0; v < 4; ++v) x[v] = ok ? pattern_feature(d[m][v], is[v]) : 0.0;
        put(m, x);
    }
    y[base + c] = !ok ? -1.0f : (dec[c] == 0xFu ? 1.0f : 0.0f);
}

// Logits of n feature rows, 16 rows per wave.
__global__ __launch_bounds__(256) void infer_kernel(Net net, const uint16_t* __restrict__ X, uint64_t n,
                                                    float* __restrict__ z) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint64_t row = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16 + (lane & 15);
    const bool valid = row < n;
    v4s xb[3];
#pragma unroll
    for (int kc = 0; kc < 3; ++kc) {
        uint16_t f[4] = {0, 0, 0, 0};
        if (valid)
            for (int j = 0; j < 4; ++j) f[j] = X[row * kFeat + kc * 16 + 4 * g + j];
        xb[kc] = pack4(f);
    }
    const float zz = mlp_wave(net, xb, lane);
    if (valid && g == 0) z[row] = zz;
}

constexpr int kTrainRows = 32;  // samples per training block (2 waves)

// Forward + backward of one block of 32 samples; writes the block's partial
// gradient of the weighted mean BCE and its partial loss.
__global__ __launch_bounds__(128) void train_kernel(Net net, const float* __restrict__ W2f,
                                                    const uint16_t* __restrict__ X, const float* __restrict__ y,
                                                    const uint32_t* __restrict__ idx, uint32_t B, float pos_w,
                                                    float neg_w, float* __restrict__ part,
                                                    float* __restrict__ loss_part) {
    __shared__ float sx[kTrainRows][kFeat];
    __shared__ float sh1[kTrainRows][kHid];
    __shared__ float spre1[kTrainRows][kHid];  // pre-activations, then their gradient
    __shared__ float spre2[kTrainRows][kHid];
    __shared__ float sdp2[kTrainRows][kHid];
    __shared__ float sdz[kTrainRows], sloss[kTrainRows];
    const int lane = threadIdx.x & 63, g = lane >> 4, w = threadIdx.x >> 6;
    const int s = w * 16 + (lane & 15);
    const uint32_t bi = blockIdx.x * kTrainRows + s;
    const bool valid = bi < B;
    const uint32_t row = valid ? idx[bi] : 0u;
    v4s xb[3];
#pragma unroll
    for (int kc = 0; kc < 3; ++kc) {
        uint16_t f[4] = {0, 0, 0, 0};
        if (valid)
            for (int j = 0; j < 4; ++j) f[j] = X[(size_t)row * kFeat + kc * 16 + 4 * g + j];
        xb[kc] = pack4(f);
        for (int j = 0; j < 4; ++j) sx[s][kc * 16 + 4 * g + j] = bf16_to_f32(f[j]);
    }
    float pre1[16], h1[16], pre2[16];
    const float z = mlp_wave(net, xb, lane, pre1, h1, pre2);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int h = nt * 16 + 4 * g + r;
            spre1[s][h] = pre1[nt * 4 + r];
            sh1[s][h] = h1[nt * 4 + r];
            spre2[s][h] = pre2[nt * 4 + r];
        }
    if (g == 0) {
        const float yy = valid ? y[row] : 0.0f;
        const float wt = valid ? (yy > 0.5f ? pos_w : neg_w) : 0.0f;
        const float p = 1.0f / (1.0f + expf(-z));
        // BCE with logits: max(z, 0) - z y + log(1 + exp(-|z|))
        sloss[s] = wt * (fmaxf(z, 0.0f) - z * yy + log1pf(expf(-fabsf(z))));
        sdz[s] = wt * (p - yy) / (float)B;
    }
    __syncthreads();
    // dpre2 = dz w3 gelu'(pre2)
    for (int e = threadIdx.x; e < kTrainRows * kHid; e += blockDim.x) {
        const int ss = e / kHid, h = e % kHid;
        sdp2[ss][h] = sdz[ss] * net.w3[h] * gelu_d(spre2[ss][h]);
    }
    __syncthreads();
    // dpre1 = (dpre2 W2^T) gelu'(pre1), W2 as the bf16 values the forward used
    for (int e = threadIdx.x; e < kTrainRows * kHid; e += blockDim.x) {
        const int ss = e / kHid, k = e % kHid;
        float acc = 0.0f;
        for (int h = 0; h < kHid; ++h) acc += sdp2[ss][h] * W2f[k * kHid + h];
        spre1[ss][k] = acc * gelu_d(spre1[ss][k]);
    }
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * kNParams;
    for (int p = threadIdx.x; p < kNParams; p += blockDim.x) {
        float acc = 0.0f;
        if (p < kOffB1) {
            const int f = p / kHid, h = p % kHid;
            for (int ss = 0; ss < kTrainRows; ++ss) acc += sx[ss][f] * spre1[ss][h];
        } else if (p < kOffW2) {
            const int h = p - kOffB1;
            for (int ss = 0; ss < kTrainRows; ++ss) acc += spre1[ss][h];
        } else if (p < kOffB2) {
            const int k = (p - kOffW2) / kHid, h = (p - kOffW2) % kHid;
            for (int ss = 0; ss < kTrainRows; ++ss) acc += sh1[ss][k] * sdp2[ss][h];
        } else if (p < kOffW3) {
            const int h = p - kOffB2;
            for (int ss = 0; ss < kTrainRows; ++ss) acc += sdp2[ss][h];
        } else if (p < kOffB3) {
            const int h = p - kOffW3;
            for (int ss = 0; ss < kTrainRows; ++ss) acc += sdz[ss] * gelu(spre2[ss][h]);
        } else {
            for (int ss = 0; ss < kTrainRows; ++ss) acc += sdz[ss];
        }
        out[p] = acc;
    }
    if (threadIdx.x == 0) {
        float l = 0.0f;
        for (int ss = 0; ss < kTrainRows; ++ss) l += sloss[ss];
        loss_part[blockIdx.x] = l;
    }
}

// grad[p] = sum over blocks in block order; loss likewise (one extra thread)
__global__ void reduce_kernel(const float* __restrict__ part, const float* __restrict__ loss_part, int nblk,
                              float* __restrict__ grad, float* __restrict__ loss, float inv_B) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNParams) {
        float acc = 0.0f;
        for (int b = 0; b < nblk; ++b) acc += part[(size_t)b * kNParams + p];
        grad[p] = acc;
    } else if (p == kNParams) {
        float acc = 0.0f;
        for (int b = 0; b < nblk; ++b) acc += loss_part[b];
        *loss = acc * inv_B;
    }
}

__global__ void adam_kernel(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParams) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    theta[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// master FP32 parameters -> MFMA operand layouts (bf16) and FP32 copies
__global__ void pack_kernel(const float* __restrict__ theta, v8s* __restrict__ W1p, v8s* __restrict__ W2p,
                            float* __restrict__ b1, float* __restrict__ b2, float* __restrict__ w3,
                            float* __restrict__ b3, float* __restrict__ W2f) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    // element j of lane l's operand for product block kk pairs with the
    // input row 32 kk + 4g + j (j < 4) or 32 kk + 16 + 4g + j - 4 (j >= 4),
    // g = l / 16, which is where mlp_wave puts those inputs; rows 48..63 of
    // the first layer are zero padding
    if (e < kW1pVec) {  // W1p[(nt*2+kk)*64 + l][j] = W1[row][nt*16 + l%16]
        const int l = e & 63, blk = e >> 6, nt = blk >> 1, kk = blk & 1, gg = l >> 4;
        v8s x;
        for (int j = 0; j < 8; ++j) {
            const int row = 32 * kk + (j < 4 ? 4 * gg + j : 16 + 4 * gg + j - 4);
            x[j] = row < kFeat ? (short)f32_to_bf16(theta[kOffW1 + row * kHid + nt * 16 + (l & 15)]) : (short)0;
        }
        W1p[e] = x;
    }
    if (e < kW2pVec) {
        const int l = e & 63, blk = e >> 6, nt = blk >> 1, kk = blk & 1, gg = l >> 4;
        v8s x;
        for (int j = 0; j < 8; ++j) {
            const int row = 32 * kk + (j < 4 ? 4 * gg + j : 16 + 4 * gg + j - 4);
            x[j] = (short)f32_to_bf16(theta[kOffW2 + row * kHid + nt * 16 + (l & 15)]);
        }
        W2p[e] = x;
    }
    if (e < kHid * kHid) W2f[e] = bf16_to_f32(f32_to_bf16(theta[kOffW2 + e]));
    if (e < kHid) {
        b1[e] = theta[kOffB1 + e];
        b2[e] = theta[kOffB2 + e];
        w3[e] = theta[kOffW3 + e];
    }
    if (e == 0) *b3 = theta[kOffB3];
}

// batch indices for training step `step`: entries of the sample list chosen
// by a counter-based hash of (seed, step, i)
__global__ void batch_kernel(const uint32_t* __restrict__ pool, uint64_t n_pool, uint64_t seed, uint64_t step,
                             uint32_t B, uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + step * 0xBF58476D1CE4E5B9ull + i * 0x94D049BB133111EBull;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    idx[i] = pool[x % n_pool];
}

}  // namespace

size_t net_bytes() {
    return (size_t)(kW1pVec + kW2pVec) * sizeof(v8s) + (3 * kHid + 1) * sizeof(float) +
           (size_t)kHid * kHid * sizeof(float);
}

ind::Net net_view(void* buf, double threshold, const float** W2f) {
    char* p = static_cast<char*>(buf);
    ind::Net n{};
    n.W1p = p;
    p += kW1pVec * sizeof(v8s);
    n.W2p = p;
    p += kW2pVec * sizeof(v8s);
    n.b1 = reinterpret_cast<const float*>(p);
    n.b2 = n.b1 + kHid;
    n.w3 = n.b2 + kHid;
    n.b3 = n.w3 + kHid;
    if (W2f) *W2f = n.b3 + 1;
    n.logit_threshold = threshold >= 1.0 ? INFINITY : (float)std::log(threshold / (1.0 - threshold));
    return n;
}

hipError_t net_pack(const float* theta, void* buf, hipStream_t stream) {
    const float* W2f = nullptr;
    const ind::Net n = net_view(buf, 0.5, &W2f);
    hipLaunchKernelGGL(pack_kernel, dim3((kHid * kHid + 255) / 256), dim3(256), 0, stream, theta,
                       static_cast<v8s*>(const_cast<void*>(n.W1p)), static_cast<v8s*>(const_cast<void*>(n.W2p)),
                       const_cast<float*>(n.b1),
                       const_cast<float*>(n.b2), const_cast<float*>(n.w3), const_cast<float*>(n.b3),
                       const_cast<float*>(W2f));
    return hipGetLastError();
}

hipError_t ind_features(const Geom& g, const Teno& t, const double* U, const uint8_t* dec, double gamma,
                        uint16_t* X, float* y, uint64_t base, hipStream_t stream) {
    if (t.M - 1 < kMembers - 1) return hipErrorInvalidValue;  // the central stencil has >= 12 cells
    hipLaunchKernelGGL(features_kernel, dim3((g.N + 255) / 256), dim3(256), 0, stream, g, t, U, dec, gamma, X, y,
                       base);
    return hipGetLastError();
}

hipError_t ind_infer(const ind::Net& net, const uint16_t* X, uint64_t n, float* z, hipStream_t stream) {
    const uint64_t waves = (n + 15) / 16;
    hipLaunchKernelGGL(infer_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, stream, net, X, n, z);
    return hipGetLastError();
}

int ind_train_blocks(uint32_t B) { return (int)((B + kTrainRows - 1) / kTrainRows); }

hipError_t ind_grad(const ind::Net& net, const float* W2f, const uint16_t* X, const float* y, const uint32_t* idx,
                    uint32_t B, float pos_w, float neg_w, float* part, float* loss_part, float* grad, float* loss,
                    hipStream_t stream) {
    const int nblk = ind_train_blocks(B);
    hipLaunchKernelGGL(train_kernel, dim3(nblk), dim3(128), 0, stream, net, W2f, X, y, idx, B, pos_w, neg_w, part,
                       loss_part);
    hipLaunchKernelGGL(reduce_kernel, dim3((kNParams + 1 + 255) / 256), dim3(256), 0, stream, part, loss_part, nblk,
                       grad, loss, 1.0f / (float)B);
    return hipGetLastError();
}

hipError_t ind_adam(float* theta, float* m, float* v, const float* grad, float lr, int step, hipStream_t stream) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, stream, theta, m, v, grad, lr, b1,
                       b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t ind_batch(const uint32_t* pool, uint64_t n_pool, uint64_t seed, uint64_t step, uint32_t B,
                     uint32_t* idx, hipStream_t stream) {
    hipLaunchKernelGGL(batch_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, pool, n_pool, seed, step, B, idx);
    return hipGetLastError();
}

}  // namespace mlr::dev
