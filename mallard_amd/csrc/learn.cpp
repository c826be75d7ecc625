#include "learn.hpp"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace mlr {

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// device allocation released on scope exit
struct DBuf {
    void* p = nullptr;
    explicit DBuf(size_t bytes) { hip_ok(hipMalloc(&p, bytes ? bytes : 8), "hipMalloc"); }
    ~DBuf() { if (p) (void)hipFree(p); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { hip_ok(hipStreamCreate(&s), "hipStreamCreate"); }
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

// network buffer packed from FP32 master parameters already on the device
struct DeviceNet {
    DBuf buf{dev::net_bytes()};
    dev::ind::Net net{};
    const float* W2f = nullptr;
    void pack(const float* theta_dev, hipStream_t s) {
        hip_ok(dev::net_pack(theta_dev, buf.p, s), "net_pack");
        net = dev::net_view(buf.p, 0.5, &W2f);
    }
};

}  // namespace

Dataset::Dataset(uint64_t capacity) : cap(capacity) {
    if (capacity == 0) fail("dataset capacity must be positive");
    hip_ok(hipMalloc(&X, capacity * kIndFeatures * sizeof(uint16_t)), "hipMalloc");
    if (hipMalloc(&y, capacity * sizeof(float)) != hipSuccess) {
        (void)hipFree(X);
        X = nullptr;
        fail("hipMalloc failed for the dataset labels");
    }
}

Dataset::~Dataset() {
    if (X) (void)hipFree(X);
    if (y) (void)hipFree(y);
}

void Dataset::get(uint64_t start, uint64_t count, uint16_t* X_out, float* y_out) const {
    if (start + count > n) fail("dataset range out of bounds");
    if (X_out)
        hip_ok(hipMemcpy(X_out, X + start * kIndFeatures, count * kIndFeatures * sizeof(uint16_t),
                         hipMemcpyDeviceToHost), "hipMemcpy");
    if (y_out) hip_ok(hipMemcpy(y_out, y + start, count * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

void mlp_forward(const float* params, const uint16_t* X, uint64_t n, float* z) {
    if (n == 0) return;
    Stream st;
    DBuf theta(kIndParams * sizeof(float)), dX(n * kIndFeatures * sizeof(uint16_t)), dz(n * sizeof(float));
    hip_ok(hipMemcpy(theta.p, params, kIndParams * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    hip_ok(hipMemcpy(dX.p, X, n * kIndFeatures * sizeof(uint16_t), hipMemcpyHostToDevice), "hipMemcpy");
    DeviceNet net;
    net.pack(theta.as<float>(), st.s);
    hip_ok(dev::ind_infer(net.net, dX.as<uint16_t>(), n, dz.as<float>(), st.s), "ind_infer");
    hip_ok(hipStreamSynchronize(st.s), "sync");
    hip_ok(hipMemcpy(z, dz.p, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

void mlp_grad(const float* params, const Dataset& ds, const uint32_t* idx, uint32_t B, float pos_w, float neg_w,
              float* grad, float* loss) {
    if (B == 0) fail("empty batch");
    for (uint32_t i = 0; i < B; ++i)
        if (idx[i] >= ds.n) fail("batch row out of range");
    Stream st;
    const int nblk = dev::ind_train_blocks(B);
    DBuf theta(kIndParams * sizeof(float)), didx(B * sizeof(uint32_t));
    DBuf part((size_t)nblk * kIndParams * sizeof(float)), lpart(nblk * sizeof(float));
    DBuf dgrad(kIndParams * sizeof(float)), dloss(sizeof(float));
    hip_ok(hipMemcpy(theta.p, params, kIndParams * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    hip_ok(hipMemcpy(didx.p, idx, B * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
    DeviceNet net;
    net.pack(theta.as<float>(), st.s);
    hip_ok(dev::ind_grad(net.net, net.W2f, ds.X, ds.y, didx.as<uint32_t>(), B, pos_w, neg_w, part.as<float>(),
                         lpart.as<float>(), dgrad.as<float>(), dloss.as<float>(), st.s), "ind_grad");
    hip_ok(hipStreamSynchronize(st.s), "sync");
    hip_ok(hipMemcpy(grad, dgrad.p, kIndParams * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    hip_ok(hipMemcpy(loss, dloss.p, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

void mlp_train(float* params, const Dataset& ds, const uint32_t* pool, uint64_t n_pool, int steps, uint32_t B,
               float lr, uint64_t seed, float pos_w, float neg_w, float* loss_hist) {
    if (n_pool == 0 || B == 0 || steps <= 0) fail("training needs rows, a batch size and steps");
    for (uint64_t i = 0; i < n_pool; ++i)
        if (pool[i] >= ds.n) fail("training row out of range");
    Stream st;
    const int nblk = dev::ind_train_blocks(B);
    DBuf theta(kIndParams * sizeof(float)), m(kIndParams * sizeof(float)), v(kIndParams * sizeof(float));
    DBuf dpool(n_pool * sizeof(uint32_t)), didx(B * sizeof(uint32_t));
    DBuf part((size_t)nblk * kIndParams * sizeof(float)), lpart(nblk * sizeof(float));
    DBuf dgrad(kIndParams * sizeof(float)), dloss((size_t)steps * sizeof(float));
    hip_ok(hipMemcpy(theta.p, params, kIndParams * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    hip_ok(hipMemcpy(dpool.p, pool, n_pool * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
    hip_ok(hipMemsetAsync(m.p, 0, kIndParams * sizeof(float), st.s), "memset");
    hip_ok(hipMemsetAsync(v.p, 0, kIndParams * sizeof(float), st.s), "memset");
    DeviceNet net;
    for (int k = 0; k < steps; ++k) {
        net.pack(theta.as<float>(), st.s);
        hip_ok(dev::ind_batch(dpool.as<uint32_t>(), n_pool, seed, (uint64_t)k, B, didx.as<uint32_t>(), st.s),
               "ind_batch");
        hip_ok(dev::ind_grad(net.net, net.W2f, ds.X, ds.y, didx.as<uint32_t>(), B, pos_w, neg_w,
                             part.as<float>(), lpart.as<float>(), dgrad.as<float>(), dloss.as<float>() + k, st.s),
               "ind_grad");
        hip_ok(dev::ind_adam(theta.as<float>(), m.as<float>(), v.as<float>(), dgrad.as<float>(), lr, k + 1, st.s),
               "ind_adam");
    }
    hip_ok(hipStreamSynchronize(st.s), "sync");
    hip_ok(hipMemcpy(params, theta.p, kIndParams * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    if (loss_hist)
        hip_ok(hipMemcpy(loss_hist, dloss.p, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

}  // namespace mlr
