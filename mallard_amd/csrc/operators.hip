// Point-wise and vector operators behind the reference's strategy
// interfaces, for callers that keep their own solver loop:
//   - RiemannSolver::calc_flux (riemann_solver.h:42-94) over a batch of face
//     points, with the device Riemann code the residual kernels use;
//   - the vector updates of TimeIntegrator::take_step
//     (time_integrator.cpp:57-163): Kokkos::deep_copy, KokkosBlas::axpy
//     (y <- y + a x) and KokkosBlas::axpby (y <- a x + b y), one element per
//     thread in the same floating-point form.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "common.hpp"
#include "device.hpp"
#include "riemann.hpp"

namespace mlr::dev {

namespace {

// One face point per thread: nrm[i][2], L/R[i][6] = (rho, u, v, p, gamma, h),
// F[i][4] = calc_flux(F, n, rho_l, u_l, p_l, gamma_l, h_l, rho_r, ...).
template <int RS>
__global__ __launch_bounds__(256) void riemann_points_kernel(int guard, int64_t n, const double* __restrict__ nrm,
                                                             const double* __restrict__ L,
                                                             const double* __restrict__ R, double* __restrict__ F) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* l = L + 6 * i;
    const double* r = R + 6 * i;
    const double ul[2] = {l[1], l[2]}, ur[2] = {r[1], r[2]};
    double f[4];
    riemann_flux(RS, guard, f, nrm[2 * i], nrm[2 * i + 1], l[0], ul, l[3], l[4], l[5], r[0], ur, r[3], r[4], r[5]);
#pragma unroll
    for (int k = 0; k < 4; ++k) F[4 * i + k] = f[k];
}

__global__ __launch_bounds__(256) void vec_update_kernel(int kind, int64_t n, double a, const double* __restrict__ x,
                                                         double b, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (kind == 0) y[i] = x[i];
    else if (kind == 1) y[i] = y[i] + a * x[i];
    else y[i] = a * x[i] + b * y[i];
}

}  // namespace

hipError_t riemann_points(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R,
                          double* F, hipStream_t stream) {
    if (n < 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks), block(256);
    switch (type) {
    case 0: hipLaunchKernelGGL((riemann_points_kernel<0>), grid, block, 0, stream, guard, n, nrm, L, R, F); break;
    case 1: hipLaunchKernelGGL((riemann_points_kernel<1>), grid, block, 0, stream, guard, n, nrm, L, R, F); break;
    case 2: hipLaunchKernelGGL((riemann_points_kernel<2>), grid, block, 0, stream, guard, n, nrm, L, R, F); break;
    case 3: hipLaunchKernelGGL((riemann_points_kernel<3>), grid, block, 0, stream, guard, n, nrm, L, R, F); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t vec_update(int kind, int64_t n, double a, const double* x, double b, double* y, hipStream_t stream) {
    if (n < 0 || kind < 0 || kind > 2) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(vec_update_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, kind, n, a, x, b, y);
    return hipGetLastError();
}

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Per-point doubles of a staged batch: normals (2), left (6), right (6),
// fluxes (4), stored as four consecutive arrays for n points.
constexpr int64_t kPointDoubles = 18;

void check_points(int type, int64_t n, const double* nrm, const double* L, const double* R, const double* F) {
    if (type < 0 || type > 3) fail("Unknown Riemann solver type.");
    if (n < 0) fail("riemann flux: negative point count");
    if (n > 0 && (!nrm || !L || !R || !F)) fail("riemann flux: NULL array");
}

// Copies the batch into dev (4 arrays of n points), evaluates it on the null
// stream and copies the fluxes back.
void run_points(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R, double* F,
                double* dev) {
    double* dn = dev;
    double* dl = dn + 2 * n;
    double* dr = dl + 6 * n;
    double* df = dr + 6 * n;
    hip_check(hipMemcpy(dn, nrm, 16 * (size_t)n, hipMemcpyHostToDevice), "riemann flux: copy in");
    hip_check(hipMemcpy(dl, L, 48 * (size_t)n, hipMemcpyHostToDevice), "riemann flux: copy in");
    hip_check(hipMemcpy(dr, R, 48 * (size_t)n, hipMemcpyHostToDevice), "riemann flux: copy in");
    hip_check(riemann_points(type, guard, n, dn, dl, dr, df, nullptr), "riemann flux: launch");
    hip_check(hipMemcpy(F, df, 32 * (size_t)n, hipMemcpyDeviceToHost), "riemann flux: copy out");
}

}  // namespace

namespace {

// Fixed staging area of the allocation-free Riemann entry point: room for
// kNoAllocPoints face points (18 doubles each, 9.4 MB per device), held by
// every device the library runs on until the process ends.
constexpr int64_t kNoAllocPoints = int64_t(1) << 16;
__device__ __align__(256) double scratch[kNoAllocPoints * kPointDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current. In host code the array's name is only a
// stand-in: kernels and copies take this address instead.
double* scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "riemann flux: device query");
    if (dev < 0 || dev >= kMaxDevices) fail("riemann flux: device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "riemann flux: scratch address");
        if (!p) fail("riemann flux: scratch address is NULL");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

int64_t riemann_flux_noalloc_capacity() { return kNoAllocPoints; }

void riemann_flux_noalloc(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R,
                          double* F) {
    check_points(type, n, nrm, L, R, F);
    if (n > kNoAllocPoints)
        fail("riemann flux (no allocation): " + std::to_string(n) + " points exceed the fixed scratch of " +
             std::to_string(kNoAllocPoints));
    double* dev = scratch_address();
    if (n == 0) return;
    run_points(type, guard, n, nrm, L, R, F, dev);
}

void riemann_flux_host(int type, int guard, int64_t n, const double* nrm, const double* L, const double* R,
                       double* F) {
    check_points(type, n, nrm, L, R, F);
    if (n == 0) return;
    struct Buf {
        void* p = nullptr;
        ~Buf() { if (p) (void)hipFree(p); }
    } buf;
    hip_check(hipMalloc(&buf.p, sizeof(double) * kPointDoubles * (size_t)n), "riemann flux: allocation");
    run_points(type, guard, n, nrm, L, R, F, static_cast<double*>(buf.p));
}

}  // namespace mlr::dev
