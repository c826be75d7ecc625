// Host side of the learned troubled-cell indicator: a device-resident
// training set (bf16 features + TENO labels) and the training / inference
// drivers over the kernels of indicator.hip.
#pragma once

#include <cstdint>

namespace mlr {

constexpr int kIndFeatures = 48;
constexpr int kIndParams = 7361;  // W1[48][64], b1[64], W2[64][64], b2[64], w3[64], b3

// Rows of (48 bf16 features, label) on the device, appended by
// Solver::collect_features; rows are never reordered.
struct Dataset {
    uint16_t* X = nullptr;  // [cap][48]
    float* y = nullptr;     // [cap]
    uint64_t n = 0, cap = 0;
    explicit Dataset(uint64_t capacity);
    ~Dataset();
    Dataset(const Dataset&) = delete;
    Dataset& operator=(const Dataset&) = delete;
    void get(uint64_t start, uint64_t count, uint16_t* X_out, float* y_out) const;
};

// Logits of n host feature rows (bf16) under host FP32 master parameters.
void mlp_forward(const float* params, const uint16_t* X, uint64_t n, float* z);

// Gradient (kIndParams) and loss of the weighted mean BCE over rows idx[0..B)
// of the set: label-1 rows weigh pos_w, label-0 rows neg_w.
void mlp_grad(const float* params, const Dataset& ds, const uint32_t* idx, uint32_t B, float pos_w, float neg_w,
              float* grad, float* loss);

// Adam on the GPU: `steps` mini-batches of B rows drawn from pool[0..n_pool)
// by a counter-based hash of (seed, step, i); params updated in place; the
// loss of every step goes to loss_hist (may be NULL).
void mlp_train(float* params, const Dataset& ds, const uint32_t* pool, uint64_t n_pool, int steps, uint32_t B,
               float lr, uint64_t seed, float pos_w, float neg_w, float* loss_hist);

}  // namespace mlr
