// TENO pseudo-inverses on the GPU (face_reconstruction.cpp:477-741): one
// thread per (cell, stencil) builds the stencil's least-squares matrix from
// the members' Legendre-basis averages in the target cell's frame, takes the
// R factor of its Householder QR and solves R^T R X = A^T, writing A+ (its
// non-zero (K-1) x (M-1) block) and the members' transformed areas straight
// into the blocked device tables. Every operation is the host precompute's,
// in the same order (no contraction), so the tables are bit-identical to the
// host's; the stencil search itself stays on the host.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace mlr::dev {

namespace {

// basis.h:81-185 Legendre polynomials P_0..P_9 as (|c|, power, sign) terms
// evaluated term by term and combined left to right (the host's evaluator)
struct LegTerm {
    double c;
    int e;
};
struct LegPoly {
    double scale;
    int nt;
    LegTerm t[5];
};
__constant__ LegPoly kLeg[10] = {
    {1.0, 1, {{1.0, 0}}},
    {1.0, 1, {{1.0, 1}}},
    {0.5, 2, {{3.0, 2}, {-1.0, 0}}},
    {0.5, 2, {{5.0, 3}, {-3.0, 1}}},
    {0.125, 3, {{35.0, 4}, {-30.0, 2}, {3.0, 0}}},
    {0.125, 3, {{63.0, 5}, {-70.0, 3}, {15.0, 1}}},
    {0.0625, 4, {{231.0, 6}, {-315.0, 4}, {105.0, 2}, {-5.0, 0}}},
    {0.0625, 4, {{429.0, 7}, {-693.0, 5}, {315.0, 3}, {-35.0, 1}}},
    {0.0078125, 5, {{6435.0, 8}, {-12012.0, 6}, {6930.0, 4}, {-1260.0, 2}, {35.0, 0}}},
    {0.0078125, 5, {{12155.0, 9}, {-25740.0, 7}, {18018.0, 5}, {-4620.0, 3}, {315.0, 1}}},
};

__device__ inline double legendre(int p, double x) {
    const LegPoly& P = kLeg[p];
    double acc = 0.0;
    for (int j = 0; j < P.nt; j++) {
        const int e = P.t[j].e;
        double term = fabs(P.t[j].c);
        for (int k = 0; k < e; k++) term *= x;
        const bool neg = P.t[j].c < 0;
        if (j == 0) acc = neg ? -term : term;
        else acc = neg ? acc - term : acc + term;
    }
    return P.scale * acc;
}

__device__ inline void gemv2(const double* A, const double* x, double* y) {
    const double t0 = A[0] * x[0] + A[1] * x[1];
    const double t1 = A[2] * x[0] + A[3] * x[1];
    y[0] = t0;
    y[1] = t1;
}

__device__ inline double tri_area(const double* v0, const double* v1, const double* v2) {
    return 0.5 * fabs(v0[0] * (v1[1] - v2[1]) + v1[0] * (v2[1] - v0[1]) + v2[0] * (v0[1] - v1[1]));
}

__device__ inline void load_tri(const PreInput& in, uint32_t c, double* v0, double* v1, double* v2) {
    const uint32_t a = in.tri[3 * (size_t)c], b = in.tri[3 * (size_t)c + 1], d = in.tri[3 * (size_t)c + 2];
    v0[0] = in.xy[2 * (size_t)a]; v0[1] = in.xy[2 * (size_t)a + 1];
    v1[0] = in.xy[2 * (size_t)b]; v1[1] = in.xy[2 * (size_t)b + 1];
    v2[0] = in.xy[2 * (size_t)d]; v2[1] = in.xy[2 * (size_t)d + 1];
}

// TenoCtx::member_row: basis averages (times the transformed area) of member
// nb in the frame (w0, Ji), and that area
template <int K>
__device__ inline void member_row(const PreInput& in, uint32_t nb, const double* w0, const double* Ji, double* row,
                                  double& at) {
    double v0[2], v1[2], v2[2], Jn[4];
    load_tri(in, nb, v0, v1, v2);
    Jn[0] = v1[0] - v0[0];
    Jn[1] = v2[0] - v0[0];
    Jn[2] = v1[1] - v0[1];
    Jn[3] = v2[1] - v0[1];
    double t0[2], t1[2], t2[2];
    for (int i = 0; i < 2; i++) { t0[i] = v0[i] - w0[i]; t1[i] = v1[i] - w0[i]; t2[i] = v2[i] - w0[i]; }
    gemv2(Ji, t0, t0);
    gemv2(Ji, t1, t1);
    gemv2(Ji, t2, t2);
    at = tri_area(t0, t1, t2);
    double qx[8], qy[8];
    for (int q = 0; q < in.nq; ++q) {
        double x[2] = {in.qp[2 * q], in.qp[2 * q + 1]};
        gemv2(Jn, x, x);
        for (int i = 0; i < 2; i++) x[i] += v0[i];
        for (int i = 0; i < 2; i++) x[i] -= w0[i];
        gemv2(Ji, x, x);
        qx[q] = x[0];
        qy[q] = x[1];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double a = 0.0;
        for (int q = 0; q < in.nq; ++q)
            a += in.qw[q] * (legendre(in.poly[2 * k], qx[q]) * legendre(in.poly[2 * k + 1], qy[q]));
        row[k] = a * at;
    }
}

// thread = (block of 64 cells, cell in block, stencil s): the order of the
// blocked tables, so consecutive lanes write consecutive doubles
template <int K, int MP>
__global__ __launch_bounds__(256) void aplus_kernel(PreInput in, double* __restrict__ A, double* __restrict__ ta,
                                                    int* __restrict__ err) {
    constexpr int KP = K - 1;
    const uint32_t blk = blockIdx.x, s = threadIdx.x & 3u, cl = threadIdx.x >> 2;
    const uint32_t c = blk * 64u + cl;
    if (c >= in.N) return;
    if (!((in.info[c] >> (20 + s)) & 1u)) return;

    // target frame
    double w0[2], w1[2], w2[2], J[4], Ji[4];
    load_tri(in, c, w0, w1, w2);
    J[0] = w1[0] - w0[0];
    J[1] = w2[0] - w0[0];
    J[2] = w1[1] - w0[1];
    J[3] = w2[1] - w0[1];
    const double det = J[0] * J[3] - J[1] * J[2];
    if (det == 0.0) {
        atomicOr(err, 1);
        return;
    }
    const double id = 1.0 / det;
    Ji[0] = J[3] * id;
    Ji[1] = -J[1] * id;
    Ji[2] = -J[2] * id;
    Ji[3] = J[0] * id;

    // least-squares matrix without its first row and column (B), areas
    double B[MP][KP];
    double row[K], at;
    bool col0 = true;
    member_row<K>(in, c, w0, Ji, row, at);
    if (row[0] - at * in.ipt[0] > 1.0e-12) col0 = false;
    for (int i = 0; i < MP; ++i) {
        const uint32_t nb = (uint32_t)in.ids[teno_mem_off(blk, MP, i, cl, (int)s)];
        member_row<K>(in, nb, w0, Ji, row, at);
        if (row[0] - at * in.ipt[0] > 1.0e-12) col0 = false;
        for (int k = 1; k < K; ++k) B[i][k - 1] = row[k] - at * in.ipt[k];
        ta[teno_mem_off(blk, MP, i, cl, (int)s)] = at;
    }
    if (!col0) {
        atomicOr(err, 2);
        return;
    }

    // R of the Householder QR (common_math.h:355-422, the host's qr_r)
    double R[MP][KP];
    for (int i = 0; i < MP; ++i)
        for (int k = 0; k < KP; ++k) R[i][k] = B[i][k];
    for (int j = 0; j < KP; j++) {
        double nx = 0.0;
        for (int i = j; i < MP; i++) nx += R[i][j] * R[i][j];
        nx = sqrt(nx);
        if (nx < 1.0e-15) continue;
        const double sgn = (R[j][j] >= 0.0) ? 1.0 : -1.0;
        const double alpha = -sgn * nx;
        double v[MP];
        double nu = 0.0;
        const int mj = MP - j;
        for (int k = 0; k < mj; k++) {
            v[k] = R[j + k][j];
            if (k == 0) v[k] -= alpha;
            nu += v[k] * v[k];
        }
        nu = sqrt(nu);
        for (int k = 0; k < mj; k++) v[k] /= nu;
        for (int c2 = 0; c2 < KP; c2++) {
            double col[MP];
            for (int a = 0; a < mj; a++) {
                double sum = 0.0;
                for (int b = 0; b < mj; b++) sum += ((a == b ? 1.0 : 0.0) - 2.0 * v[a] * v[b]) * R[j + b][c2];
                col[a] = sum;
            }
            for (int a = 0; a < mj; a++) R[j + a][c2] = col[a];
        }
        for (int i = 0; i < j; i++)
            for (int c2 = 0; c2 < KP; c2++) R[i][c2] = R[i][c2] + 0.0;
    }

    // forward_sub(R, B, Y, kp, kp, mp, tL, tB): R^T Y = B^T
    double Y[KP][MP];
    for (int i = 0; i < KP; i++)
        for (int jj = 0; jj < MP; jj++) {
            double sum = 0.0;
            for (int k = 0; k < i; k++) sum += R[k][i] * Y[k][jj];
            Y[i][jj] = (B[jj][i] - sum) / R[i][i];
        }
    // back_sub(R, Y, X, kp, kp, mp): R X = Y
    double X[KP][MP];
    for (int i = KP - 1; i >= 0; i--)
        for (int jj = 0; jj < MP; jj++) {
            double sum = 0.0;
            for (int k = i + 1; k < KP; k++) sum += R[i][k] * X[k][jj];
            X[i][jj] = (Y[i][jj] - sum) / R[i][i];
        }
    for (int i = 0; i < MP; ++i)
        for (int k = 0; k < KP; ++k) A[teno_a_off(blk, MP, KP, i, k, cl, (int)s)] = X[k][i];
}

// thread = (block of 64 cells, cell in block, face point fq), the order of
// the operator's table; the member loop writes lin(fq, i) for every member
// of the cell's central stencil
__global__ __launch_bounds__(256) void linear_operator_kernel(Teno t, const uint32_t* __restrict__ info, uint32_t N,
                                                              int FQ, uint64_t total, double* __restrict__ lin) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int fq = (int)(gid % (uint64_t)FQ);
    const uint64_t rest = gid / (uint64_t)FQ;
    const uint32_t cl = (uint32_t)(rest & 63u);
    const uint64_t blk = rest >> 6;
    const uint32_t c = (uint32_t)(blk * 64 + cl);
    const int KP = t.K - 1, Mp = t.M - 1;
    const bool on = c < N && ((info[c] >> 20) & 1u);
    // face-point basis values: the cell's table, or with compact tables the
    // row of the face's reference edge (areas folded into A)
    const double* B = t.fbasis + ((blk * FQ + fq) * KP) * 64 + cl;
    size_t bs = 64;
    if (t.compact) {  // one face point per face: fq is the slot
        const uint32_t e = c < N ? (info[c] >> (24 + 2 * fq)) & 3u : 0u;
        B = t.fbc + (size_t)e * KP;
        bs = 1;
    }
    for (int i = 0; i < Mp; ++i) {
        double w = 0.0;
        if (on) {  // central stencil (s = 0), member i
            const double* Ai = t.A + teno_a_off(blk, Mp, KP, i, 0, cl, 0);
            double acc = 0.0;
            for (int k = 0; k < KP; ++k) acc = acc + B[k * bs] * Ai[(size_t)k * 256];
            w = t.compact ? acc : acc * t.ta[teno_mem_off(blk, Mp, i, cl, 0)];
        }
        lin[teno_lin_off(blk, Mp, i, cl, FQ, fq)] = w;
    }
}

// one thread per member-table entry e = (block, member) * 256 + cell * 4 + stencil;
// its A+ entries are those of the same (block, member) and slot, 256 apart over k
__global__ __launch_bounds__(256) void fold_areas_kernel(double* __restrict__ A, const double* __restrict__ ta,
                                                         uint64_t total, int KP) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const double t = ta[gid];
    double* a = A + (gid >> 8) * (uint64_t)KP * 256 + (gid & 255u);
    for (int k = 0; k < KP; ++k) a[(uint64_t)k * 256] = a[(uint64_t)k * 256] * t;
}

}  // namespace

hipError_t teno_linear_operator(const Teno& t, const uint32_t* info, uint32_t N, int FQ, double* lin,
                                hipStream_t stream) {
    if (FQ < 1 || FQ > 8 || t.M < 2) return hipErrorInvalidValue;
    const uint64_t total = (uint64_t)((N + 63) / 64) * 64 * FQ;
    hipLaunchKernelGGL(linear_operator_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, t, info,
                       N, FQ, total, lin);
    return hipGetLastError();
}

hipError_t teno_fold_areas(double* A, const double* ta, uint64_t rows, int KP, hipStream_t stream) {
    if (KP < 1) return hipErrorInvalidValue;
    const uint64_t total = rows * 64;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(fold_areas_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, A, ta, total,
                       KP);
    return hipGetLastError();
}

bool aplus_supported(int K, int M) { return (K == 3 && M == 6) || (K == 6 && M == 12) || (K == 10 && M == 20); }

hipError_t teno_aplus(const PreInput& in, int K, int M, double* A, double* ta, int* err, hipStream_t stream) {
    const dim3 grid((in.N + 63) / 64), block(256);
    if (in.nq < 1 || in.nq > 8) return hipErrorInvalidValue;
    if (K == 3 && M == 6) hipLaunchKernelGGL((aplus_kernel<3, 5>), grid, block, 0, stream, in, A, ta, err);
    else if (K == 6 && M == 12) hipLaunchKernelGGL((aplus_kernel<6, 11>), grid, block, 0, stream, in, A, ta, err);
    else if (K == 10 && M == 20) hipLaunchKernelGGL((aplus_kernel<10, 19>), grid, block, 0, stream, in, A, ta, err);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace mlr::dev
