/*
 * CPU oracle: clean-room FP64 restatement of Mallard's explicit residual
 * evaluation, SSP-RK stage update and CFL time step (reference numbering,
 * AoS Kokkos LayoutRight layout, atomics in the multi-threaded mode).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, as the reference the HIP path is compared
 * with and as the timed CPU baseline.
 *
 * With one thread the residual is accumulated in the canonical order of a
 * one-thread reference run (interior faces in ascending id, then each boundary
 * zone in input order, each zone in generation order; mesh.cpp:690-733,
 * solver_rhs.cpp:44-55, flux_functor.h:156-161). With more threads the
 * face contributions are added with atomics, like the reference.
 *
 * Parity pins (see tests/): Riemann fluxes, EOS constants, QR / substitutions /
 * gemm / inverse / triangle helpers and the constant-RHS integrator test are
 * pinned by the reference's own unit-test values. Mesh generation, stencils,
 * pseudo-inverses, OI, the TENO kernel, boundary fluxes, dt and whole
 * trajectories have no reference-held outputs; they are pinned by closed
 * forms, textbook formulas and brute-force restatements: counts, face-index
 * formulas, volume sums, closed normals, free-stream preservation, A+ A = I
 * on the fitted subspace, SVD pseudo-inverse agreement, linear-field offsets,
 * conservation, SSPRK3 amplification, per-face spectral radii and dt on
 * unjittered meshes, the ENO branch on a straight jump (both weightings,
 * raw weights and the +ipt/ta0 offset included), the directional stencils
 * against a brute-force Type-4 search, wall / fixed-state / p_out residuals,
 * Toro's HLL flux, and the Roe flux (no reference counterpart) at exact
 * Rankine-Hugoniot shocks; each pin fails on planted mistakes
 * (tests/test_oracle_mutations.py). Also parity unpinned beyond the
 * invariants above: the order of
 * equal-distance ring members (libstdc++ std::sort ties,
 * face_reconstruction.cpp:253-257).
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int g_vacuum_guard = 0;  // set per residual evaluation (CORRECTED weighting only)
thread_local std::string g_err;

constexpr int ND = 2;   // common_typedef.h:19
constexpr int NC = 4;   // common_typedef.h:20
constexpr int NP = 5;   // common_typedef.h:21

// ---------------------------------------------------------------- math
// common_math.h:31-53 (dot<2>: a0*b0 + a1*b1)
inline double dot2(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1]; }
// common_math.h:71-74
inline double norm2_2(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1]); }
// common_math.h:101-106: multiply by the reciprocal norm
inline void unit2(const double* v, double* u) {
    const double inv = 1 / norm2_2(v);
    u[0] = v[0] * inv;
    u[1] = v[1] * inv;
}
// common_math.h:145-159
void invert2(const double* A, double* Ai) {
    const double det = A[0] * A[3] - A[1] * A[2];
    if (det == 0.0) throw std::runtime_error("Matrix is singular.");
    const double id = 1.0 / det;
    Ai[0] = A[3] * id;
    Ai[1] = -A[1] * id;
    Ai[2] = -A[2] * id;
    Ai[3] = A[0] * id;
}
// common_math.h:162-194
void invert3(const double* A, double* Ai) {
    const double c11 = A[4] * A[8] - A[5] * A[7];
    const double c12 = -(A[3] * A[8] - A[5] * A[6]);
    const double c13 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c11 + A[1] * c12 + A[2] * c13;
    if (det == 0.0) throw std::runtime_error("Matrix is singular.");
    const double id = 1.0 / det;
    const double c21 = -(A[1] * A[8] - A[2] * A[7]);
    const double c22 = A[0] * A[8] - A[2] * A[6];
    const double c23 = -(A[0] * A[7] - A[1] * A[6]);
    const double c31 = A[1] * A[5] - A[2] * A[4];
    const double c32 = -(A[0] * A[5] - A[2] * A[3]);
    const double c33 = A[0] * A[4] - A[1] * A[3];
    Ai[0] = c11 * id; Ai[1] = c21 * id; Ai[2] = c31 * id;
    Ai[3] = c12 * id; Ai[4] = c22 * id; Ai[5] = c32 * id;
    Ai[6] = c13 * id; Ai[7] = c23 * id; Ai[8] = c33 * id;
}
// common_math.h:207-213 (in-place safe)
inline void gemv2(const double* A, const double* x, double* y) {
    const double t0 = A[0] * x[0] + A[1] * x[1];
    const double t1 = A[2] * x[0] + A[3] * x[1];
    y[0] = t0;
    y[1] = t1;
}
// common_math.h:240-345: sequential k accumulation from 0.0, in-place safe
void gemm(const double* A, const double* B, double* C, int m, int n, int p, int q,
          bool tA, bool tB) {
    std::vector<double> tmp;
    double* Co = C;
    if (C == A || C == B) { tmp.assign((size_t)m * q, 0.0); Co = tmp.data(); }
    if (!tA && !tB) {
        if (n != p) throw std::runtime_error("Invalid matrix dimensions.");
        for (int i = 0; i < m; i++)
            for (int j = 0; j < q; j++) {
                double s = 0.0;
                for (int k = 0; k < n; k++) s += A[i * n + k] * B[k * q + j];
                Co[i * q + j] = s;
            }
    } else if (tA && !tB) {
        if (m != p) throw std::runtime_error("Invalid matrix dimensions.");
        for (int i = 0; i < n; i++)
            for (int j = 0; j < q; j++) {
                double s = 0.0;
                for (int k = 0; k < m; k++) s += A[k * n + i] * B[k * q + j];
                Co[i * q + j] = s;
            }
    } else if (!tA && tB) {
        if (n != q) throw std::runtime_error("Invalid matrix dimensions.");
        for (int i = 0; i < m; i++)
            for (int j = 0; j < p; j++) {
                double s = 0.0;
                for (int k = 0; k < n; k++) s += A[i * n + k] * B[j * q + k];
                Co[i * p + j] = s;
            }
    } else {
        if (m != q) throw std::runtime_error("Invalid matrix dimensions.");
        for (int i = 0; i < n; i++)
            for (int j = 0; j < p; j++) {
                double s = 0.0;
                for (int k = 0; k < m; k++) s += A[k * n + i] * B[j * q + k];
                Co[i * p + j] = s;
            }
    }
    if (Co != C) for (int i = 0; i < m * q; i++) C[i] = Co[i];
}
// common_math.h:355-422: R of a Householder QR, reflectors applied as explicit
// dense Q_j products through gemm (the rounding depends on that).
void qr_noq(const double* A, double* R, int m, int n) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) R[i * n + j] = A[i * n + j];
    std::vector<double> v(m), Q((size_t)m * m);
    for (int j = 0; j < n; j++) {
        double nx = 0.0;
        for (int i = j; i < m; i++) nx += R[i * n + j] * R[i * n + j];
        nx = std::sqrt(nx);
        if (nx < 1.0e-15) continue;
        const double sgn = (R[j * n + j] >= 0.0) ? 1.0 : -1.0;
        const double alpha = -sgn * nx;
        double nu = 0.0;
        for (int k = 0; k < m - j; k++) {
            v[k] = R[(j + k) * n + j];
            if (k == 0) v[k] -= alpha;
            nu += v[k] * v[k];
        }
        nu = std::sqrt(nu);
        for (int k = 0; k < m - j; k++) v[k] /= nu;
        for (int i = 0; i < m; i++)
            for (int k = 0; k < m; k++) Q[(size_t)i * m + k] = (i == k) ? 1.0 : 0.0;
        for (int i = j; i < m; i++)
            for (int k = j; k < m; k++) Q[(size_t)i * m + k] -= 2.0 * v[i - j] * v[k - j];
        gemm(Q.data(), R, R, m, m, m, n, false, false);
    }
}
// common_math.h:436-457
void forward_sub(const double* L, const double* B, double* X, int m, int n, int p, bool tL,
                 bool tB) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < p; j++) {
            double s = 0.0;
            for (int k = 0; k < i; k++) {
                const int iL = tL ? k * m + i : i * n + k;
                s += L[iL] * X[k * p + j];
            }
            const int iB = tB ? j * n + i : i * p + j;
            const int iL = tL ? i * m + i : i * n + i;
            X[i * p + j] = (B[iB] - s) / L[iL];
        }
}
// common_math.h:471-493
void back_sub(const double* U, const double* B, double* X, int m, int n, int p, bool tU,
              bool tB) {
    for (int i = m - 1; i >= 0; i--)
        for (int j = 0; j < p; j++) {
            double s = 0.0;
            const int end = tU ? n : m;
            for (int k = i + 1; k < end; k++) {
                const int iU = tU ? k * m + i : i * n + k;
                s += U[iU] * X[k * p + j];
            }
            const int iB = tB ? j * n + i : i * p + j;
            const int iU = tU ? i * m + i : i * n + i;
            X[i * p + j] = (B[iB] - s) / U[iU];
        }
}
// common_math.h:510-516
inline double tri_area(const double* v0, const double* v1, const double* v2) {
    return 0.5 * std::fabs(v0[0] * (v1[1] - v2[1]) + v1[0] * (v2[1] - v0[1]) +
                           v2[0] * (v0[1] - v1[1]));
}
// common_math.h:534-542
inline void tri_J(const double* v0, const double* v1, const double* v2, double* J) {
    J[0] = v1[0] - v0[0];
    J[1] = v2[0] - v0[0];
    J[2] = v1[1] - v0[1];
    J[3] = v2[1] - v0[1];
}

// ---------------------------------------------------------------- basis
// Legendre polynomials P_0..P_9 in the expanded form of basis.h:85-94 and their
// n-th derivatives (basis.h:110-183): scale * (t_0 -/+ t_1 ...), each term
// computed as ((c * x) * x) ... left to right and the terms combined left to
// right. The coefficient products are exact small integers.
struct LegTerm { double c; int e; };
struct LegPoly { double scale; int nt; LegTerm t[5]; };
const LegPoly kLeg[10] = {
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
double legendre_d(int n, int p, double x) {
    if (p > 9 || n > 9) throw std::runtime_error("Legendre basis implemented for p <= 9.");
    if (n > p) return 0.0;
    const LegPoly& P = kLeg[p];
    double acc = 0.0;
    bool first = true;
    for (int j = 0; j < P.nt; j++) {
        const int e = P.t[j].e;
        if (e < n) continue;
        double c = std::fabs(P.t[j].c);
        for (int k = 0; k < n; k++) c *= (double)(e - k);
        double term = c;
        for (int k = 0; k < e - n; k++) term *= x;
        if (first) {
            acc = (P.t[j].c < 0) ? -term : term;
            first = false;
        } else {
            acc = (P.t[j].c < 0) ? acc - term : acc + term;
        }
    }
    return P.scale * acc;
}
// basis.h:67-78 (monomial; the derivative keeps the reference's (p-k+1)
// factors and negative powers)
double monomial(int p, double x) { return std::pow(x, (double)p); }
double monomial_d(int n, int p, double x) {
    double r = std::pow(x, (double)(p - n));
    for (int k = 0; k < n; ++k) r *= (double)(p - k + 1);
    return r;
}
double basis1(int basis, int p, double x) {
    return basis == ORC_BASIS_LEGENDRE ? legendre_d(0, p, x) : monomial(p, x);
}
double basisd(int basis, int n, int p, double x) {
    return basis == ORC_BASIS_LEGENDRE ? legendre_d(n, p, x) : monomial_d(n, p, x);
}
// basis.h:58-61
double basis2(int basis, int px, int py, double x, double y) {
    return basis1(basis, px, x) * basis1(basis, py, y);
}

// ---------------------------------------------------------------- quadrature
// Standard Gauss-Legendre (on [-1,1]) and Dunavant (reference triangle, weights
// summing to 1) rules, with the decimal constants of quadrature.cpp:23-273 so
// that the parsed doubles are identical.
struct Quad { int dim, order, n; std::vector<double> pts, w; };
Quad gauss_legendre(int order) {
    Quad q{1, order, 0, {}, {}};
    switch (order) {
    case 1: q.pts = {0.0}; q.w = {2.0}; break;
    case 2: q.pts = {-0.577350269189625764509149, 0.577350269189625764509149}; q.w = {1.0, 1.0}; break;
    case 3: q.pts = {-0.774596669241483377035853, 0.0, 0.774596669241483377035853};
            q.w = {0.55555555555555555555556, 0.88888888888888888888889, 0.55555555555555555555556}; break;
    case 4: q.pts = {-0.861136311594052575223946, -0.339981043584856264802666, 0.339981043584856264802666, 0.861136311594052575223946};
            q.w = {0.34785484513745385737306, 0.65214515486254614262694, 0.65214515486254614262694, 0.34785484513745385737306}; break;
    case 5: q.pts = {-0.906179845938663992797627, -0.538469310105683091036314, 0.0, 0.538469310105683091036314, 0.906179845938663992797627};
            q.w = {0.23692688505618908751426, 0.47862867049936646804129, 0.56888888888888888888889, 0.47862867049936646804129, 0.23692688505618908751426}; break;
    case 6: q.pts = {-0.932469514203152027812302, -0.661209386466264513661400, -0.238619186083196908630502, 0.238619186083196908630502, 0.661209386466264513661400, 0.932469514203152027812302};
            q.w = {0.17132449237917034504030, 0.36076157304813860756983, 0.46791393457269104738987, 0.46791393457269104738987, 0.36076157304813860756983, 0.17132449237917034504030}; break;
    case 7: q.pts = {-0.949107912342758524526190, -0.741531185599394439863865, -0.405845151377397166906607, 0.0, 0.405845151377397166906607, 0.741531185599394439863865, 0.949107912342758524526190};
            q.w = {0.12948496616886969327061, 0.27970539148927666790147, 0.38183005050511894495037, 0.41795918367346938775510, 0.38183005050511894495037, 0.27970539148927666790147, 0.12948496616886969327061}; break;
    default: throw std::runtime_error("Gauss-Legendre quadrature rule of order " + std::to_string(order) + " not implemented.");
    }
    q.n = (int)q.w.size();
    return q;
}
Quad dunavant(int order) {
    Quad q{2, order, 0, {}, {}};
    switch (order) {
    case 1: q.pts = {1.0 / 3.0, 1.0 / 3.0}; q.w = {1.0}; break;
    case 2: q.pts = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
            q.w = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0}; break;
    case 3: q.pts = {1.0 / 3.0, 1.0 / 3.0, 0.6, 0.2, 0.2, 0.6, 0.2, 0.2};
            q.w = {-0.5625, 0.52083333333333333333333333333333, 0.52083333333333333333333333333333, 0.52083333333333333333333333333333}; break;
    case 4: q.pts = {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.108103018168070,
                     0.445948490915965, 0.445948490915965, 0.816847572980459, 0.091576213509771,
                     0.091576213509771, 0.816847572980459, 0.091576213509771, 0.091576213509771};
            q.w = {0.223381589678011, 0.223381589678011, 0.223381589678011, 0.109951743655322, 0.109951743655322, 0.109951743655322}; break;
    case 5: q.pts = {0.333333333333333, 0.333333333333333, 0.059715871789770, 0.470142064105115,
                     0.470142064105115, 0.059715871789770, 0.470142064105115, 0.470142064105115,
                     0.797426985353087, 0.101286507323456, 0.101286507323456, 0.797426985353087,
                     0.101286507323456, 0.101286507323456};
            q.w = {0.225000000000000, 0.132394152788506, 0.132394152788506, 0.132394152788506, 0.125939180544827, 0.125939180544827, 0.125939180544827}; break;
    default: throw std::runtime_error("Triangle Dunavant quadrature rule of order " + std::to_string(order) + " not implemented.");
    }
    q.n = (int)q.w.size();
    return q;
}

}  // namespace

// ---------------------------------------------------------------- mesh
struct orc_mesh {
    int type = 0;
    uint32_t nx = 0, ny = 0;
    uint32_t n_cells = 0, n_nodes = 0, n_faces = 0, n_faces_real = 0;
    std::vector<double> node_coords, cell_coords, cell_volume, face_area, face_normals;
    std::vector<uint32_t> nodes_of_cell, off_noc, faces_of_cell, off_foc, nodes_of_face, off_nof;
    std::vector<int32_t> cells_of_face;
    struct Zone { std::string name; bool interior; std::vector<uint32_t> faces; };
    std::vector<Zone> zones;

    uint32_t n_nodes_of_cell(uint32_t c) const { return off_noc[c + 1] - off_noc[c]; }
    uint32_t n_faces_of_cell(uint32_t c) const { return off_foc[c + 1] - off_foc[c]; }
    uint32_t node_of_cell(uint32_t c, int l) const { return nodes_of_cell[off_noc[c] + l]; }
    uint32_t face_of_cell(uint32_t c, int l) const { return faces_of_cell[off_foc[c] + l]; }
    uint32_t node_of_face(uint32_t f, int l) const { return nodes_of_face[off_nof[f] + l]; }
    double nc(uint32_t n, int d) const { return node_coords[2 * (size_t)n + d]; }
    double cc(uint32_t c, int d) const { return cell_coords[2 * (size_t)c + d]; }
    int32_t cof(uint32_t f, int s) const { return cells_of_face[2 * (size_t)f + s]; }

    // mesh.cpp:127-165: the cell itself plus its face neighbours, sorted, unique
    void neighbors1(uint32_t c, std::vector<uint32_t>& out) const {
        out.push_back(c);
        for (uint32_t l = 0; l < n_faces_of_cell(c); ++l) {
            const uint32_t f = face_of_cell(c, l);
            const int32_t c0 = cof(f, 0), c1 = cof(f, 1);
            if (c1 == -1) continue;
            out.push_back(c0 == (int32_t)c ? (uint32_t)c1 : (uint32_t)c0);
        }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }

    void build_csr(const std::vector<std::vector<uint32_t>>& noc,
                   const std::vector<std::vector<uint32_t>>& foc,
                   const std::vector<std::vector<uint32_t>>& nof) {
        off_noc.assign(n_cells + 1, 0);
        off_foc.assign(n_cells + 1, 0);
        off_nof.assign(n_faces + 1, 0);
        nodes_of_cell.clear(); faces_of_cell.clear(); nodes_of_face.clear();
        for (uint32_t c = 0; c < n_cells; ++c) {
            off_noc[c + 1] = off_noc[c] + (uint32_t)noc[c].size();
            off_foc[c + 1] = off_foc[c] + (uint32_t)foc[c].size();
            for (auto v : noc[c]) nodes_of_cell.push_back(v);
            for (auto v : foc[c]) faces_of_cell.push_back(v);
        }
        for (uint32_t f = 0; f < n_faces; ++f) {
            off_nof[f + 1] = off_nof[f] + (uint32_t)nof[f].size();
            for (auto v : nof[f]) nodes_of_face.push_back(v);
        }
    }

    // mesh.cpp:218-227
    void face_areas() {
        face_area.assign(n_faces, 0.0);
        for (uint32_t f = 0; f < n_faces; ++f) {
            const uint32_t a = node_of_face(f, 0), b = node_of_face(f, 1);
            const double dx = nc(b, 0) - nc(a, 0), dy = nc(b, 1) - nc(a, 1);
            face_area[f] = std::sqrt(dx * dx + dy * dy);
        }
    }
    // mesh.cpp:181-216
    void cell_volumes() {
        cell_volume.assign(n_cells, 0.0);
        for (uint32_t c = 0; c < n_cells; ++c) {
            const uint32_t k = n_nodes_of_cell(c);
            double v0[2], v1[2], v2[2], v3[2];
            auto ld = [&](int l, double* v) { v[0] = nc(node_of_cell(c, l), 0); v[1] = nc(node_of_cell(c, l), 1); };
            if (k == 3) {
                ld(0, v0); ld(1, v1); ld(2, v2);
                cell_volume[c] = tri_area(v0, v1, v2);
            } else if (k == 4) {
                ld(0, v0); ld(1, v1); ld(2, v2); ld(3, v3);
                const double a1 = tri_area(v0, v1, v2), a2 = tri_area(v0, v2, v3);
                cell_volume[c] = a1 + a2;
            } else {
                throw std::runtime_error("Unknown cell type.");
            }
        }
    }
    // mesh.cpp:167-179
    void cell_centroids() {
        cell_coords.assign(2 * (size_t)n_cells, 0.0);
        for (uint32_t c = 0; c < n_cells; ++c) {
            double x = 0.0, y = 0.0;
            const uint32_t k = n_nodes_of_cell(c);
            for (uint32_t l = 0; l < k; ++l) {
                x += nc(node_of_cell(c, l), 0);
                y += nc(node_of_cell(c, l), 1);
            }
            cell_coords[2 * (size_t)c] = x / k;
            cell_coords[2 * (size_t)c + 1] = y / k;
        }
    }
    // mesh.cpp:229-261 (phantom faces: 0/0 -> NaN normals, as in the reference)
    void face_normals_() {
        face_normals.assign(2 * (size_t)n_faces, 0.0);
        for (uint32_t f = 0; f < n_faces; ++f) {
            const uint32_t a = node_of_face(f, 0), b = node_of_face(f, 1);
            const double x0 = nc(a, 0), y0 = nc(a, 1), x1 = nc(b, 0), y1 = nc(b, 1);
            const double dx = x1 - x0, dy = y1 - y0;
            const double mag = std::sqrt(dx * dx + dy * dy);
            double n0 = dy / mag * face_area[f];
            double n1 = -dx / mag * face_area[f];
            const int32_t c0 = cof(f, 0);
            const double xc = cc(c0, 0), yc = cc(c0, 1);
            const double xf = 0.5 * (x0 + x1), yf = 0.5 * (y0 + y1);
            const double d = (xf - xc) * n0 + (yf - yc) * n1;
            if (d < 0) { n0 *= -1; n1 *= -1; }
            face_normals[2 * (size_t)f] = n0;
            face_normals[2 * (size_t)f + 1] = n1;
        }
    }
};

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Build-specific jitter extension (no reference counterpart): interior node
// (i, j) moves by jitter * (dx, dy) * (2u - 1) with u from a counter-based
// splitmix64 stream keyed by (seed, node id).
void jitter_nodes(orc_mesh& m, uint32_t nx, uint32_t ny, double dx, double dy, double jitter,
                  uint64_t seed) {
    if (jitter == 0.0) return;
    const uint64_t key = splitmix64(seed);
    for (uint32_t i = 1; i < nx; ++i)
        for (uint32_t j = 1; j < ny; ++j) {
            const uint64_t id = (uint64_t)i * (ny + 1) + j;
            const double u1 = (double)(splitmix64(key + 2 * id) >> 11) * 0x1.0p-53;
            const double u2 = (double)(splitmix64(key + 2 * id + 1) >> 11) * 0x1.0p-53;
            m.node_coords[2 * id] += jitter * dx * (2.0 * u1 - 1.0);
            m.node_coords[2 * id + 1] += jitter * dy * (2.0 * u2 - 1.0);
        }
}

void finish_zones(orc_mesh& m, std::vector<uint32_t>& fi, std::vector<uint32_t>& fr,
                  std::vector<uint32_t>& ft, std::vector<uint32_t>& fl, std::vector<uint32_t>& fb) {
    std::sort(fi.begin(), fi.end());
    fi.erase(std::unique(fi.begin(), fi.end()), fi.end());
    m.zones.clear();
    m.zones.push_back({"interior", true, fi});
    m.zones.push_back({"right", false, fr});
    m.zones.push_back({"top", false, ft});
    m.zones.push_back({"left", false, fl});
    m.zones.push_back({"bottom", false, fb});
}

// mesh.cpp:305-557
void init_cart(orc_mesh& m, uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter,
               uint64_t seed) {
    m.nx = nx; m.ny = ny;
    m.n_cells = nx * ny;
    m.n_nodes = (nx + 1) * (ny + 1);
    m.n_faces = 2 * m.n_cells + nx + ny;
    m.n_faces_real = m.n_faces;
    m.node_coords.assign(2 * (size_t)m.n_nodes, 0.0);
    m.cells_of_face.assign(2 * (size_t)m.n_faces, 0);
    std::vector<std::vector<uint32_t>> noc(m.n_cells, std::vector<uint32_t>(4)),
        foc(m.n_cells, std::vector<uint32_t>(4)), nof(m.n_faces, std::vector<uint32_t>(2));
    const double dx = Lx / nx, dy = Ly / ny;
    for (uint32_t i = 0; i < nx + 1; ++i)
        for (uint32_t j = 0; j < ny + 1; ++j) {
            const uint32_t n = i * (ny + 1) + j;
            m.node_coords[2 * (size_t)n] = i * dx;
            m.node_coords[2 * (size_t)n + 1] = j * dy;
        }
    jitter_nodes(m, nx, ny, dx, dy, jitter, seed);
    for (uint32_t c = 0; c < m.n_cells; ++c) {
        const uint32_t ic = c / ny, jc = c % ny;
        const uint32_t n_tr = (ic + 1) * (ny + 1) + jc + 1, n_tl = ic * (ny + 1) + jc + 1;
        const uint32_t n_bl = ic * (ny + 1) + jc, n_br = (ic + 1) * (ny + 1) + jc;
        const uint32_t f_r = (2 * ny + 1) * (ic + 1) + jc, f_t = (2 * ny + 1) * ic + jc + ny + 1;
        const uint32_t f_l = (2 * ny + 1) * ic + jc, f_b = (2 * ny + 1) * ic + jc + ny;
        noc[c] = {n_tr, n_tl, n_bl, n_br};
        foc[c] = {f_r, f_t, f_l, f_b};
        auto set = [&](uint32_t f, int32_t a, int32_t b) { m.cells_of_face[2 * (size_t)f] = a; m.cells_of_face[2 * (size_t)f + 1] = b; };
        set(f_r, c, ic == nx - 1 ? -1 : (int32_t)(c + ny));
        set(f_t, c, jc == ny - 1 ? -1 : (int32_t)(c + 1));
        set(f_l, c, ic == 0 ? -1 : (int32_t)(c - ny));
        set(f_b, c, jc == 0 ? -1 : (int32_t)(c - 1));
        nof[f_r] = {n_br, n_tr}; nof[f_t] = {n_tr, n_tl};
        nof[f_l] = {n_tl, n_bl}; nof[f_b] = {n_bl, n_br};
    }
    std::vector<uint32_t> fi, fr, ft, fl, fb;
    for (uint32_t c = 0; c < m.n_cells; ++c) {
        const uint32_t ic = c / ny, jc = c % ny;
        (ic == nx - 1 ? fr : fi).push_back(foc[c][0]);
        (jc == ny - 1 ? ft : fi).push_back(foc[c][1]);
        (ic == 0 ? fl : fi).push_back(foc[c][2]);
        (jc == 0 ? fb : fi).push_back(foc[c][3]);
    }
    finish_zones(m, fi, fr, ft, fl, fb);
    m.build_csr(noc, foc, nof);
    m.face_areas();
    m.cell_volumes();
    m.cell_centroids();
    m.face_normals_();
}

// mesh.cpp:559-826 (n_faces over-allocated and the node order of each
// triangle as in the reference)
void init_cart_tri(orc_mesh& m, uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter,
                   uint64_t seed) {
    m.nx = nx; m.ny = ny;
    m.n_cells = 2 * nx * ny;
    m.n_nodes = (nx + 1) * (ny + 1);
    m.n_faces = 3 * m.n_cells + nx + ny;
    m.n_faces_real = 3 * nx * ny + nx + ny;
    m.node_coords.assign(2 * (size_t)m.n_nodes, 0.0);
    m.cells_of_face.assign(2 * (size_t)m.n_faces, 0);
    std::vector<std::vector<uint32_t>> noc(m.n_cells, std::vector<uint32_t>(3)),
        foc(m.n_cells, std::vector<uint32_t>(3)), nof(m.n_faces, std::vector<uint32_t>(2));
    const double dx = Lx / nx, dy = Ly / ny;
    for (uint32_t i = 0; i < nx + 1; ++i)
        for (uint32_t j = 0; j < ny + 1; ++j) {
            const uint32_t n = i * (ny + 1) + j;
            m.node_coords[2 * (size_t)n] = i * dx;
            m.node_coords[2 * (size_t)n + 1] = j * dy;
        }
    jitter_nodes(m, nx, ny, dx, dy, jitter, seed);
    for (uint32_t q = 0; q < nx * ny; ++q) {
        const uint32_t ic = q / ny, jc = q % ny;
        const uint32_t cr = 2 * q, cl = cr + 1;
        const uint32_t n_tr = (ic + 1) * (ny + 1) + jc + 1, n_tl = ic * (ny + 1) + jc + 1;
        const uint32_t n_bl = ic * (ny + 1) + jc, n_br = (ic + 1) * (ny + 1) + jc;
        const uint32_t f_r = (3 * ny + 1) * (ic + 1) + jc;
        const uint32_t f_t = (3 * ny + 1) * ic + 2 * jc + ny + 2;
        const uint32_t f_l = (3 * ny + 1) * ic + jc;
        const uint32_t f_b = (3 * ny + 1) * ic + 2 * jc + ny;
        const uint32_t f_d = (3 * ny + 1) * ic + 2 * jc + ny + 1;
        noc[cr] = {n_br, n_tr, n_bl};
        noc[cl] = {n_tl, n_bl, n_tr};
        foc[cr] = {f_b, f_r, f_d};
        foc[cl] = {f_t, f_l, f_d};
        auto set = [&](uint32_t f, int32_t a, int32_t b) { m.cells_of_face[2 * (size_t)f] = a; m.cells_of_face[2 * (size_t)f + 1] = b; };
        set(f_r, cr, ic == nx - 1 ? -1 : (int32_t)(cl + 2 * ny));
        set(f_t, cl, jc == ny - 1 ? -1 : (int32_t)(cl + 1));
        set(f_l, cl, ic == 0 ? -1 : (int32_t)(cr - 2 * ny));
        set(f_b, cr, jc == 0 ? -1 : (int32_t)(cr - 1));
        set(f_d, cr, cl);
        nof[f_r] = {n_br, n_tr}; nof[f_t] = {n_tr, n_tl}; nof[f_l] = {n_tl, n_bl};
        nof[f_b] = {n_bl, n_br}; nof[f_d] = {n_bl, n_tr};
    }
    std::vector<uint32_t> fi, fr, ft, fl, fb;
    for (uint32_t q = 0; q < nx * ny; ++q) {
        const uint32_t ic = q / ny, jc = q % ny;
        const uint32_t cr = 2 * q, cl = cr + 1;
        (ic == nx - 1 ? fr : fi).push_back(foc[cr][1]);
        (jc == ny - 1 ? ft : fi).push_back(foc[cl][0]);
        (ic == 0 ? fl : fi).push_back(foc[cl][1]);
        (jc == 0 ? fb : fi).push_back(foc[cr][0]);
        fi.push_back(foc[cr][2]);
    }
    finish_zones(m, fi, fr, ft, fl, fb);
    m.build_csr(noc, foc, nof);
    m.face_areas();
    m.cell_volumes();
    m.cell_centroids();
    m.face_normals_();
}

// mesh.cpp:828-848 (8-degree ramp for x > 0.5)
void init_wedge(orc_mesh& m, uint32_t nx, uint32_t ny, double Lx, double Ly) {
    init_cart(m, nx, ny, Lx, Ly, 0.0, 0);
    const double th = 8 * M_PI / 180.0, xw = 0.5;
    for (uint32_t n = 0; n < m.n_nodes; ++n) {
        const double x = m.node_coords[2 * (size_t)n], y = m.node_coords[2 * (size_t)n + 1];
        if (x > xw) {
            const double yb = (x - xw) * std::tan(th);
            m.node_coords[2 * (size_t)n + 1] = (y / Ly) * (Ly - yb) + yb;
        }
    }
    m.face_areas();
    m.cell_volumes();
    m.cell_centroids();
    m.face_normals_();
}

// ---------------------------------------------------------------- physics
// physics.cpp:69-73, physics.h:819-867
struct Euler {
    double gamma, p_ref, T_ref, rho_ref, p_min, p_max, R, cp, cv;
    void set(double g, double pr, double Tr, double rr, double pmin, double pmax) {
        gamma = g; p_ref = pr; T_ref = Tr; rho_ref = rr; p_min = pmin; p_max = pmax;
        R = p_ref / (T_ref * rho_ref);
        cp = R * gamma / (gamma - 1.0);
        cv = cp / gamma;
    }
    double p_from_rho_e(double rho, double e) const {
        return std::fmax(p_min, std::fmin(p_max, (gamma - 1.0) * rho * e));
    }
    double sos(double p, double rho) const { return std::sqrt(gamma * p / rho); }
    void prims(double* W, const double* U) const {
        const double rho = U[0];
        const double u[2] = {U[1] / rho, U[2] / rho};
        const double E = U[3] / rho;
        const double e = E - 0.5 * dot2(u, u);
        const double p = p_from_rho_e(rho, e);
        const double T = e / cv;
        const double h = e + p / rho;
        W[0] = u[0]; W[1] = u[1]; W[2] = p; W[3] = T; W[4] = h;
    }
};

// ---------------------------------------------------------------- Riemann
// riemann_solver.h:205-240 (averaged form)
void pvrs(const double* Wl, const double* Wr, double& ps, double& rls, double& rrs) {
    const double rl = Wl[0], ul = Wl[1], pl = Wl[2], gl = Wl[3];
    const double rr = Wr[0], ur = Wr[1], pr = Wr[2], gr = Wr[3];
    const double al = std::sqrt(gl * pl / rl), ar = std::sqrt(gr * pr / rr);
    const double ravg = 0.5 * (rl + rr), aavg = 0.5 * (al + ar);
    const double us = 0.5 * (ul + ur) + 0.5 * (pl - pr) / (ravg * aavg);
    ps = 0.5 * (pl + pr) + 0.5 * (ul - ur) * ravg * aavg;
    rls = rl + (ul - us) * ravg / aavg;
    rrs = rr - (ur - us) * ravg / aavg;
}
// riemann_solver.h:242-271
void trrs(const double* Wl, const double* Wr, double& ps, double& rls, double& rrs) {
    const double rl = Wl[0], ul = Wl[1], pl = Wl[2], gl = Wl[3];
    const double rr = Wr[0], ur = Wr[1], pr = Wr[2], gr = Wr[3];
    const double al = std::sqrt(gl * pl / rl), ar = std::sqrt(gr * pr / rr);
    const double zl = (gl - 1.0) / (2.0 * gl), zr = (gr - 1.0) / (2.0 * gr);
    const double Plr = std::pow((pl / pr), zl);
    const double us = (Plr * ul / al + ur / ar + 2.0 * (1.0 - Plr) / (gl - 1.0)) / (Plr / al + 1.0 / ar);
    double bl = 1.0 + (gl - 1.0) / (2.0 * al) * (ul - us);
    double br = 1.0 + (gr - 1.0) / (2.0 * ar) * (us - ur);
    if (g_vacuum_guard) { bl = std::fmax(bl, 0.0); br = std::fmax(br, 0.0); }
    ps = 0.5 * (pl * std::pow(bl, (1.0 / zl)) + pr * std::pow(br, (1.0 / zr)));
    rls = rl * std::pow(ps / pl, 1.0 / gl);
    rrs = rr * std::pow(ps / pr, 1.0 / gr);
}
// riemann_solver.h:273-306 (p_star enters from PVRS)
void tsrs(const double* Wl, const double* Wr, double& ps, double& rls, double& rrs) {
    const double rl = Wl[0], ul = Wl[1], pl = Wl[2], gl = Wl[3];
    const double rr = Wr[0], ur = Wr[1], pr = Wr[2], gr = Wr[3];
    const double gml = (gl - 1.0) / (gl + 1.0), gmr = (gr - 1.0) / (gr + 1.0);
    const double Al = 2.0 / (gl + 1.0) / rl, Bl = gml * pl;
    const double Ar = 2.0 / (gr + 1.0) / rr, Br = gmr * pr;
    double p0 = ps;
    p0 = std::fmax(0.0, p0);
    const double gl0 = std::sqrt(Al / (p0 + Bl)), gr0 = std::sqrt(Ar / (p0 + Br));
    ps = (gl0 * pl + gr0 * pr - (ur - ul)) / (gl0 + gr0);
    rls = rl * (ps / pl + gml) / (gml * ps / pl + 1.0);
    rrs = rr * (ps / pr + gmr) / (gmr * ps / pr + 1.0);
}
// riemann_solver.h:308-329
void anrs(const double* Wl, const double* Wr, double& ps, double& rls, double& rrs) {
    const double pl = Wl[2], pr = Wr[2];
    const double pmax = std::fmax(pl, pr), pmin = std::fmin(pl, pr);
    const double qmax = pmax / pmin, quser = 2.0;
    pvrs(Wl, Wr, ps, rls, rrs);
    if ((qmax < quser) && (pmin <= ps) && (ps <= pmax)) {
    } else if (ps < pmin) {
        trrs(Wl, Wr, ps, rls, rrs);
    } else {
        tsrs(Wl, Wr, ps, rls, rrs);
    }
}

struct LR {
    double uln, urn, al, ar, Ul[4], Ur[4], Fl[4], Fr[4];
};
// riemann_solver.h:338-370 (shared preliminaries of all three solvers)
inline void prelim(LR& s, const double* n, double rl, const double* ul, double pl, double gl,
                   double hl, double rr, const double* ur, double pr, double gr, double hr) {
    s.uln = dot2(ul, n);
    s.urn = dot2(ur, n);
    const double ul2 = dot2(ul, ul), ur2 = dot2(ur, ur);
    s.al = std::sqrt(gl * pl / rl);
    s.ar = std::sqrt(gr * pr / rr);
    const double El = (hl + 0.5 * ul2) * rl - pl, Er = (hr + 0.5 * ur2) * rr - pr;
    s.Ul[0] = rl; s.Ul[1] = rl * ul[0]; s.Ul[2] = rl * ul[1]; s.Ul[3] = El;
    s.Ur[0] = rr; s.Ur[1] = rr * ur[0]; s.Ur[2] = rr * ur[1]; s.Ur[3] = Er;
    s.Fl[0] = rl * s.uln;
    s.Fl[1] = rl * ul[0] * s.uln + pl * n[0];
    s.Fl[2] = rl * ul[1] * s.uln + pl * n[1];
    s.Fl[3] = (El + pl) * s.uln;
    s.Fr[0] = rr * s.urn;
    s.Fr[1] = rr * ur[0] * s.urn + pr * n[0];
    s.Fr[2] = rr * ur[1] * s.urn + pr * n[1];
    s.Fr[3] = (Er + pr) * s.urn;
}

void flux_rusanov(double* F, const double* n, double rl, const double* ul, double pl, double gl,
                  double hl, double rr, const double* ur, double pr, double gr, double hr) {
    LR s;
    prelim(s, n, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
    const double Smax = std::fmax(std::fabs(s.uln) + s.al, std::fabs(s.urn) + s.ar);
    for (int i = 0; i < NC; i++) F[i] = 0.5 * (s.Fl[i] + s.Fr[i] + Smax * (s.Ul[i] - s.Ur[i]));
}
// riemann_solver.h:377-439
void flux_hll(double* F, const double* n, double rl, const double* ul, double pl, double gl,
              double hl, double rr, const double* ur, double pr, double gr, double hr) {
    LR s;
    prelim(s, n, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
    const double Wl[4] = {rl, s.uln, pl, gl}, Wr[4] = {rr, s.urn, pr, gr};
    double ps, rls, rrs;
    anrs(Wl, Wr, ps, rls, rrs);
    const double ql = (ps <= pl) ? 1.0 : std::sqrt(1.0 + (gl + 1.0) / (2.0 * gl) * (ps / pl - 1.0));
    const double qr = (ps <= pr) ? 1.0 : std::sqrt(1.0 + (gr + 1.0) / (2.0 * gr) * (ps / pr - 1.0));
    const double Sl = s.uln - s.al * ql, Sr = s.urn + s.ar * qr;
    if (0.0 <= Sl) for (int i = 0; i < NC; i++) F[i] = s.Fl[i];
    else if (Sr <= 0.0) for (int i = 0; i < NC; i++) F[i] = s.Fr[i];
    else for (int i = 0; i < NC; i++)
        F[i] = (Sr * s.Fl[i] - Sl * s.Fr[i] + Sl * Sr * (s.Ur[i] - s.Ul[i])) / (Sr - Sl);
}
// riemann_solver.h:441-519 (HLLC "variant 2")
void flux_hllc(double* F, const double* n, double rl, const double* ul, double pl, double gl,
               double hl, double rr, const double* ur, double pr, double gr, double hr) {
    LR s;
    prelim(s, n, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
    const double Wl[4] = {rl, s.uln, pl, gl}, Wr[4] = {rr, s.urn, pr, gr};
    double ps, rls, rrs;
    anrs(Wl, Wr, ps, rls, rrs);
    const double ql = (ps <= pl) ? 1.0 : std::sqrt(1.0 + (gl + 1.0) / (2.0 * gl) * (ps / pl - 1.0));
    const double qr = (ps <= pr) ? 1.0 : std::sqrt(1.0 + (gr + 1.0) / (2.0 * gr) * (ps / pr - 1.0));
    const double Sl = s.uln - s.al * ql, Sr = s.urn + s.ar * qr;
    const double Ss = (pr - pl + rl * s.uln * (Sl - s.uln) - rr * s.urn * (Sr - s.urn)) /
                      (rl * (Sl - s.uln) - rr * (Sr - s.urn));
    if (0.0 <= Sl) {
        for (int i = 0; i < NC; i++) F[i] = s.Fl[i];
    } else if (Sr <= 0.0) {
        for (int i = 0; i < NC; i++) F[i] = s.Fr[i];
    } else {
        const double Ds[4] = {0.0, n[0], n[1], Ss};
        const double Plr = 0.5 * (pl + pr + rl * (Sl - s.uln) * (Ss - s.uln) + rr * (Sr - s.urn) * (Ss - s.urn));
        if (Ss >= 0.0) {
            for (int i = 0; i < NC; i++) F[i] = (Ss * (Sl * s.Ul[i] - s.Fl[i]) + Sl * Plr * Ds[i]) / (Sl - Ss);
        } else {
            for (int i = 0; i < NC; i++) F[i] = (Ss * (Sr * s.Ur[i] - s.Fr[i]) + Sr * Plr * Ds[i]) / (Sr - Ss);
        }
    }
}
// Roe flux with Harten's entropy fix (textbook form, Toro ch. 11). No
// reference counterpart; pinned by its exact resolution of isolated shocks.
void flux_roe(double* F, const double* n, double rl, const double* ul, double pl, double gl,
              double hl, double rr, const double* ur, double pr, double gr, double hr) {
    LR s;
    prelim(s, n, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
    const double g = 0.5 * (gl + gr);
    const double Hl = hl + 0.5 * dot2(ul, ul), Hr = hr + 0.5 * dot2(ur, ur);
    const double sl = std::sqrt(rl), sr = std::sqrt(rr), isum = 1.0 / (sl + sr);
    const double u = (sl * ul[0] + sr * ur[0]) * isum;
    const double v = (sl * ul[1] + sr * ur[1]) * isum;
    const double H = (sl * Hl + sr * Hr) * isum;
    const double q2 = u * u + v * v;
    const double a2 = (g - 1.0) * (H - 0.5 * q2);
    const double a = std::sqrt(std::fmax(a2, 1e-300));
    const double un = u * n[0] + v * n[1];
    const double ut = -u * n[1] + v * n[0];
    const double rho = sl * sr;
    const double dr = rr - rl, dp = pr - pl;
    const double dun = s.urn - s.uln;
    const double dut = (-ur[0] * n[1] + ur[1] * n[0]) - (-ul[0] * n[1] + ul[1] * n[0]);
    const double al1 = (dp - rho * a * dun) / (2.0 * a2);
    const double al2 = dr - dp / a2;
    const double al3 = rho * dut;
    const double al4 = (dp + rho * a * dun) / (2.0 * a2);
    auto fix = [&](double lam) {
        const double d = 0.1 * a;
        const double al = std::fabs(lam);
        return al < d ? (lam * lam + d * d) / (2.0 * d) : al;
    };
    const double l1 = fix(un - a), l2 = std::fabs(un), l3 = std::fabs(un), l4 = fix(un + a);
    const double r1[4] = {1.0, u - a * n[0], v - a * n[1], H - a * un};
    const double r2[4] = {1.0, u, v, 0.5 * q2};
    const double r3[4] = {0.0, -n[1], n[0], ut};
    const double r4[4] = {1.0, u + a * n[0], v + a * n[1], H + a * un};
    for (int i = 0; i < NC; i++)
        F[i] = 0.5 * (s.Fl[i] + s.Fr[i]) -
               0.5 * (l1 * al1 * r1[i] + l2 * al2 * r2[i] + l3 * al3 * r3[i] + l4 * al4 * r4[i]);
}

typedef void (*FluxFn)(double*, const double*, double, const double*, double, double, double,
                       double, const double*, double, double, double);
FluxFn flux_fn(int t) {
    switch (t) {
    case ORC_RS_RUSANOV: return flux_rusanov;
    case ORC_RS_HLL: return flux_hll;
    case ORC_RS_HLLC: return flux_hllc;
    case ORC_RS_ROE: return flux_roe;
    default: throw std::runtime_error("Unknown Riemann solver type.");
    }
}

}  // namespace

// ---------------------------------------------------------------- solver
// time_integrator.cpp:57-163: FE, RK4 and SSPRK3 as the sequence of KokkosBlas
// axpy / axpby updates the reference issues (y <- y + a x; y <- a x + b y).
void blas_axpy(int nt, double a, const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < (int64_t)y.size(); ++i) y[i] = y[i] + a * x[i];
}
void blas_axpby(int nt, double a, const std::vector<double>& x, double b, std::vector<double>& y) {
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < (int64_t)y.size(); ++i) y[i] = a * x[i] + b * y[i];
}

template <class RHS>
void integrate_step(int type, double dt, int nt, std::vector<double>& U, std::vector<double>* rhs,
                    std::vector<double>& Utmp, RHS calc) {
    auto axpy = [nt](double a, const std::vector<double>& x, std::vector<double>& y) { blas_axpy(nt, a, x, y); };
    auto axpby = [nt](double a, const std::vector<double>& x, double b, std::vector<double>& y) {
        blas_axpby(nt, a, x, b, y);
    };
    if (type == ORC_TI_FE) {
        calc(U.data(), rhs[0].data());
        axpy(dt, rhs[0], U);
    } else if (type == ORC_TI_RK4) {
        const double co[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        calc(U.data(), rhs[0].data());
        Utmp = U;
        axpy(dt / 2.0, rhs[0], Utmp);
        calc(Utmp.data(), rhs[1].data());
        Utmp = U;
        axpy(dt / 2.0, rhs[1], Utmp);
        calc(Utmp.data(), rhs[2].data());
        Utmp = U;
        axpy(dt, rhs[2], Utmp);
        calc(Utmp.data(), rhs[3].data());
        for (int k = 0; k < 4; k++) axpy(dt * co[k], rhs[k], U);
    } else if (type == ORC_TI_SSPRK3) {
        const double co[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
        calc(U.data(), rhs[0].data());
        Utmp = U;
        axpy(dt, rhs[0], Utmp);
        calc(Utmp.data(), rhs[1].data());
        axpby(3.0 / 4.0, U, 1.0 / 4.0, Utmp);
        axpy(dt / 4.0, rhs[1], Utmp);
        calc(Utmp.data(), rhs[2].data());
        for (int k = 0; k < 3; k++) axpy(dt * co[k], rhs[k], U);
    } else {
        throw std::runtime_error("Unknown time integrator type.");
    }
}

struct orc_solver {
    orc_mesh* m = nullptr;
    orc_config cfg{};
    Euler phys{};
    int nthreads = 1;
    double t = 0.0;
    // TENO state (face_reconstruction.cpp:170-788)
    int n_dof = 1, M = 0;
    std::vector<uint8_t> poly;  // [n_dof][2]
    Quad qcell, qface;
    std::vector<uint32_t> off_groups, off_st, st;
    std::vector<uint64_t> off_rm;  // 64-bit offsets (the reference's uint32 ones overflow at p=3 on 8M cells)
    std::vector<double> rm, ta, ipt, OI;
    // state
    std::vector<double> U, Utmp, prim, face, cfl_local;
    std::vector<double> rhs[4];
    std::vector<uint8_t> decisions;

    int nq() const { return qface.n; }
    double& fs(uint32_t f, int q, int side, int v) {
        return face[(((size_t)f * nq() + q) * 2 + side) * NC + v];
    }

    // ------------------------------------------------------------ TENO init
    // face_reconstruction.cpp:170-215
    void calc_sizes_and_poly() {
        n_dof = 1;
        int den = 1;
        for (int i = 1; i <= ND; ++i) { den *= i; n_dof *= cfg.order + i; }
        n_dof /= den;
        M = (uint16_t)(cfg.stencil_factor * n_dof);
        poly.assign(2 * n_dof, 0);
        int idx = 0;
        int ind[2];
        for (int p = 0; p <= cfg.order; p++) {
            ind[0] = p; ind[1] = 0;
            poly[2 * idx] = ind[0]; poly[2 * idx + 1] = ind[1]; ++idx;
            while (ind[1] < p) {
                if (ind[0] > 0) { ind[0] -= 1; ind[1] += 1; }
                poly[2 * idx] = ind[0]; poly[2 * idx + 1] = ind[1]; ++idx;
            }
        }
    }
    // face_reconstruction.cpp:217-265
    void next_ring(std::vector<std::vector<uint32_t>>& rings, uint32_t tgt) const {
        std::vector<uint32_t> nr;
        for (auto c : rings.back()) {
            std::vector<uint32_t> nb;
            m->neighbors1(c, nb);
            for (auto x : nb) {
                bool old = false;
                for (auto& r : rings)
                    if (std::find(r.begin(), r.end(), x) != r.end()) { old = true; break; }
                const bool innext = std::find(nr.begin(), nr.end(), x) != nr.end();
                if (!old && !innext) nr.push_back(x);
            }
        }
        std::vector<std::pair<uint32_t, double>> d;
        for (auto c : nr) {
            double dist = 0.0;
            for (int i = 0; i < ND; i++) {
                const double dd = m->cc(c, i) - m->cc(tgt, i);
                dist += dd * dd;  // std::pow(x, 2) folds to x * x
            }
            d.push_back(std::make_pair(c, dist));
        }
        std::sort(d.begin(), d.end(), [](auto& a, auto& b) { return a.second < b.second; });
        nr.clear();
        for (auto& pr : d) nr.push_back(pr.first);
        rings.push_back(nr);
    }
    // face_reconstruction.cpp:267-286
    std::vector<uint32_t> stencil_centered(uint32_t c) const {
        std::vector<uint32_t> s{c};
        std::vector<std::vector<uint32_t>> rings{s};
        while ((int)s.size() < M) {
            next_ring(rings, c);
            for (auto x : rings.back()) {
                if ((int)s.size() == M) break;
                s.push_back(x);
            }
        }
        return s;
    }
    // face_reconstruction.cpp:288-404
    std::vector<std::vector<uint32_t>> stencils_directional(uint32_t c) const {
        const int ns = m->n_faces_of_cell(c);
        std::vector<std::vector<uint32_t>> S(ns);
        std::vector<bool> grew(ns);
        std::vector<std::vector<double>> Jinv(ns, std::vector<double>(4));
        for (int s = 0; s < ns; ++s) {
            const uint32_t f = m->face_of_cell(c, s);
            const uint32_t a = m->node_of_face(f, 0), b = m->node_of_face(f, 1);
            double v0[2], v1[2], v2[2], J[4];
            for (int i = 0; i < ND; i++) { v0[i] = m->cc(c, i); v1[i] = m->nc(a, i); v2[i] = m->nc(b, i); }
            tri_J(v0, v1, v2, J);
            invert2(J, Jinv[s].data());
        }
        for (auto& x : S) x.push_back(c);
        std::vector<std::vector<uint32_t>> rings{{c}};
        bool done = false;
        while (!done) {
            next_ring(rings, c);
            std::fill(grew.begin(), grew.end(), false);
            for (int s = 0; s < ns; ++s) {
                for (auto x : rings.back()) {
                    if ((int)S[s].size() == M) break;
                    double dx[2], dt[2];
                    for (int i = 0; i < ND; i++) dx[i] = m->cc(x, i) - m->cc(c, i);
                    gemv2(Jinv[s].data(), dx, dt);
                    bool in = true;
                    for (int i = 0; i < ND; i++) if (dt[i] < 0.0) in = false;
                    if (in) { S[s].push_back(x); grew[s] = true; }
                }
            }
            done = true;
            for (int s = 0; s < ns; ++s) {
                if (m->cof(m->face_of_cell(c, s), 1) != -1 && (int)S[s].size() < M && grew[s]) {
                    done = false;
                    break;
                }
            }
        }
        for (int s = 0; s < ns; ++s) {
            if (m->cof(m->face_of_cell(c, s), 1) == -1) { S[s].clear(); continue; }
            rings.clear();
            rings.push_back(S[s]);
            while ((int)S[s].size() < M) {
                next_ring(rings, c);
                for (auto x : rings.back()) {
                    if ((int)S[s].size() == M) break;
                    S[s].push_back(x);
                }
            }
        }
        return S;
    }
    // face_reconstruction.cpp:406-475. Each cell's stencils depend on the
    // mesh alone, so cells are searched in parallel and concatenated in cell
    // order (the same arrays as a serial search).
    void compute_stencils() {
        const int64_t nc = m->n_cells;
        std::vector<std::vector<std::vector<uint32_t>>> per((size_t)nc);
        int not_full = 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1024) reduction(| : not_full)
        for (int64_t ci = 0; ci < nc; ++ci) {
            const uint32_t c = (uint32_t)ci;
            auto& S = per[(size_t)ci];
            S.push_back(stencil_centered(c));
            for (auto& x : stencils_directional(c)) S.push_back(std::move(x));
            for (auto& st_ : S)
                if ((int)st_.size() < M && !st_.empty()) not_full = 1;
        }
        if (not_full) throw std::runtime_error("Stencil is not full.");
        off_groups.assign(1, 0);
        off_st.assign(1, 0);
        size_t total = 0;
        for (const auto& S : per)
            for (const auto& x : S) total += x.size();
        st.clear();
        st.reserve(total);
        for (auto& S : per) {
            for (auto& x : S) {
                for (auto v : x) st.push_back(v);
                off_st.push_back((uint32_t)st.size());
            }
            off_groups.push_back((uint32_t)off_st.size() - 1);
            std::vector<std::vector<uint32_t>>().swap(S);
        }
    }
    double b2(int px, int py, double x, double y) const { return basis2(cfg.basis, px, py, x, y); }
    // face_reconstruction.cpp:477-741 for one cell: its stencils' pseudo-
    // inverses into rm_out (K x size per stencil, stencils in order) and
    // transformed member areas into ta_out. Cell 0's central stencil defines
    // ipt (Q20), so cell 0 runs first with set_ipt.
    void matrices_of_cell(uint32_t c, bool set_ipt, double* rm_out, double* ta_out) {
        const int K = n_dof;
        double J[4], Ji[4], w0[2], w1[2], w2[2];
        for (int i = 0; i < ND; i++) {
            w0[i] = m->nc(m->node_of_cell(c, 0), i);
            w1[i] = m->nc(m->node_of_cell(c, 1), i);
            w2[i] = m->nc(m->node_of_cell(c, 2), i);
        }
        tri_J(w0, w1, w2, J);
        invert2(J, Ji);
        const uint32_t s0 = off_groups[c];
        const int gsz = off_groups[c + 1] - s0;
        for (int sl = 0; sl < gsz; ++sl) {
            const uint32_t si = s0 + sl;
            const uint32_t soff = off_st[si];
            const int ssz = off_st[si + 1] - soff;
            if (ssz == 0) continue;
            std::vector<double> at(ssz), A((size_t)ssz * K), R((size_t)ssz * K), Y((size_t)K * ssz);
            for (int nl = 0; nl < ssz; ++nl) {
                const uint32_t nb = st[soff + nl];
                double v0[2], v1[2], v2[2], Jn[4];
                for (int i = 0; i < ND; i++) {
                    v0[i] = m->nc(m->node_of_cell(nb, 0), i);
                    v1[i] = m->nc(m->node_of_cell(nb, 1), i);
                    v2[i] = m->nc(m->node_of_cell(nb, 2), i);
                }
                tri_J(v0, v1, v2, Jn);
                double t0[2], t1[2], t2[2];
                for (int i = 0; i < ND; i++) { t0[i] = v0[i] - w0[i]; t1[i] = v1[i] - w0[i]; t2[i] = v2[i] - w0[i]; }
                gemv2(Ji, t0, t0);
                gemv2(Ji, t1, t1);
                gemv2(Ji, t2, t2);
                at[nl] = tri_area(t0, t1, t2);
                std::vector<double> qp(2 * (size_t)qcell.n);
                for (int q = 0; q < qcell.n; ++q) {
                    double x[2] = {qcell.pts[2 * q], qcell.pts[2 * q + 1]};
                    gemv2(Jn, x, x);
                    for (int i = 0; i < ND; i++) x[i] += v0[i];
                    for (int i = 0; i < ND; i++) x[i] -= w0[i];
                    gemv2(Ji, x, x);
                    qp[2 * q] = x[0]; qp[2 * q + 1] = x[1];
                }
                for (int k = 0; k < K; ++k) {
                    const size_t ind = (size_t)nl * K + k;
                    A[ind] = 0.0;
                    for (int q = 0; q < qcell.n; ++q)
                        A[ind] += qcell.w[q] * b2(poly[2 * k], poly[2 * k + 1], qp[2 * q], qp[2 * q + 1]);
                    A[ind] *= at[nl];
                }
            }
            if (set_ipt && sl == 0)
                for (int k = 0; k < K; ++k) ipt[k] = A[k] / at[0];
            for (int nl = 0; nl < ssz; ++nl)
                for (int k = 0; k < K; ++k) A[(size_t)nl * K + k] -= at[nl] * ipt[k];
            bool col0 = true;
            for (int nl = 0; nl < ssz; ++nl)
                if (A[(size_t)nl * K] > 1.0e-12) { col0 = false; break; }  // no fabs, as the reference
            if (col0) {
                std::vector<double> B((size_t)(ssz - 1) * (K - 1));
                for (int nl = 1; nl < ssz; ++nl)
                    for (int k = 1; k < K; ++k) B[(size_t)(nl - 1) * (K - 1) + (k - 1)] = A[(size_t)nl * K + k];
                qr_noq(B.data(), R.data(), ssz - 1, K - 1);
                forward_sub(R.data(), B.data(), Y.data(), K - 1, K - 1, ssz - 1, true, true);
                back_sub(R.data(), Y.data(), B.data(), K - 1, K - 1, ssz - 1, false, false);
                for (int k = 0; k < K; ++k)
                    for (int nl = 0; nl < ssz; ++nl)
                        A[(size_t)k * ssz + nl] = (k == 0 || nl == 0) ? 0.0 : B[(size_t)(k - 1) * (ssz - 1) + (nl - 1)];
            } else {
                qr_noq(A.data(), R.data(), ssz, K);
                forward_sub(R.data(), A.data(), Y.data(), K, K, ssz, true, true);
                back_sub(R.data(), Y.data(), A.data(), K, K, ssz, false, false);
            }
            std::memcpy(rm_out + (size_t)K * (soff - off_st[s0]), A.data(), A.size() * 8);
            std::memcpy(ta_out + (soff - off_st[s0]), at.data(), (size_t)ssz * 8);
        }
    }
    // face_reconstruction.cpp:477-741: the stencil matrices and member areas
    // of every cell. Entries are laid out like the stencil ids (K per member,
    // one per member), so every cell's slice is known in advance and the cells
    // after cell 0 are computed in parallel.
    void compute_matrices() {
        const int K = n_dof;
        const uint32_t nc = m->n_cells;
        for (uint32_t c = 0; c < nc; ++c) {
            if (m->n_nodes_of_cell(c) != 3) throw std::runtime_error("TENO has only been implemented for triangular cells.");
            if (off_st[off_groups[c] + 1] == off_st[off_groups[c]])
                throw std::runtime_error("Central stencil is empty for cell " + std::to_string(c) + ".");
        }
        ipt.assign(K, 0.0);
        off_rm.resize(off_st.size());
        for (size_t i = 0; i < off_st.size(); ++i) off_rm[i] = (uint64_t)K * off_st[i];
        rm.assign((size_t)K * st.size(), 0.0);
        ta.assign(st.size(), 0.0);
        if (nc == 0) return;
        matrices_of_cell(0, true, rm.data(), ta.data());
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1024)
        for (int64_t ci = 1; ci < (int64_t)nc; ++ci) {
            const uint32_t c = (uint32_t)ci;
            const size_t o = off_st[off_groups[c]];
            matrices_of_cell(c, false, rm.data() + (size_t)K * o, ta.data() + o);
        }
    }
    // face_reconstruction.cpp:743-788
    void compute_oi() {
        const int K = n_dof;
        OI.assign((size_t)K * K, 0.0);
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) {
                double acc = 0.0;
                for (int k = 1; k < K; k++)
                    for (int q = 0; q < qcell.n; ++q) {
                        double di = 1.0, dj = 1.0;
                        for (int d = 0; d < ND; d++)
                            di *= basisd(cfg.basis, poly[2 * k + d], poly[2 * i + d], qcell.pts[2 * q + d]);
                        for (int d = 0; d < ND; d++)
                            dj *= basisd(cfg.basis, poly[2 * k + d], poly[2 * j + d], qcell.pts[2 * q + d]);
                        acc += qcell.w[q] * di * dj;
                    }
                OI[(size_t)i * K + j] = acc;
            }
    }

    void init() {
        phys.set(cfg.gamma, cfg.p_ref, cfg.T_ref, cfg.rho_ref, cfg.p_min, cfg.p_max);
        if (cfg.recon == ORC_RECON_FO) {
            qface = gauss_legendre(1);  // face_reconstruction.cpp:41-44
        } else {
            calc_sizes_and_poly();
            const int qc = cfg.quad_order_cell > 0 ? cfg.quad_order_cell : cfg.order + 1;
            const int qf = cfg.quad_order_face > 0 ? cfg.quad_order_face : (cfg.order + 1) / 2;
            qcell = dunavant(qc);
            qface = gauss_legendre(qf);
            compute_stencils();
            compute_matrices();
            compute_oi();
        }
        const size_t nc = m->n_cells;
        U.assign(nc * NC, 0.0);
        Utmp.assign(nc * NC, 0.0);
        prim.assign(nc * NP, 0.0);
        cfl_local.assign(nc, 0.0);
        for (auto& r : rhs) r.assign(nc * NC, 0.0);
        face.assign((size_t)m->n_faces * nq() * 2 * NC, 0.0);
        decisions.assign(nc * NC, 0);
    }

    // ------------------------------------------------------------ residual
    // face_reconstruction.cpp:59-99 (side 1 of boundary faces is never read)
    void fo_faces(const double* Ux) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t f = 0; f < (int64_t)m->n_faces; ++f) {
            const int32_t c0 = m->cof(f, 0), c1 = m->cof(f, 1);
            for (int v = 0; v < NC; ++v) {
                fs(f, 0, 0, v) = Ux[(size_t)c0 * NC + v];
                fs(f, 0, 1, v) = c1 >= 0 ? Ux[(size_t)c1 * NC + v] : 0.0;
            }
        }
    }
    // Positivity-preserving scaling of the reconstructed face states toward the
    // cell average (Zhang & Shu, J. Comput. Phys. 229 (2010) 8918): one factor
    // theta in [0, 1] per cell keeps rho and p >= eps at every face point.
    // No reference counterpart (only the CORRECTED variant uses it).
    double pressure_raw(const double* u) const {
        return (phys.gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
    }
    void positivity_scale(uint32_t c, const double* Ux) {
        const double* Uc = &Ux[(size_t)c * NC];
        const double pc = pressure_raw(Uc);
        const double eps = std::fmin(1.0e-13, std::fmin(Uc[0], pc));
        const int nfc = m->n_faces_of_cell(c);
        double theta = 1.0;
        for (int fl = 0; fl < nfc; fl++) {
            const uint32_t f = m->face_of_cell(c, fl);
            const int side = m->cof(f, 0) == (int32_t)c ? 0 : 1;
            for (int q = 0; q < nq(); q++) {
                double u[NC];
                for (int v = 0; v < NC; v++) u[v] = fs(f, q, side, v);
                double t = 1.0;
                if (u[0] < eps) t = (Uc[0] - eps) / (Uc[0] - u[0]);
                double w[NC];
                for (int v = 0; v < NC; v++) w[v] = Uc[v] + t * (u[v] - Uc[v]);
                if (pressure_raw(w) < eps) {
                    double lo = 0.0, hi = t;
                    for (int it = 0; it < 40; it++) {
                        const double mid = 0.5 * (lo + hi);
                        for (int v = 0; v < NC; v++) w[v] = Uc[v] + mid * (u[v] - Uc[v]);
                        if (pressure_raw(w) < eps) hi = mid; else lo = mid;
                    }
                    t = lo;
                }
                theta = std::fmin(theta, t);
            }
        }
        if (theta < 1.0) {
            for (int fl = 0; fl < nfc; fl++) {
                const uint32_t f = m->face_of_cell(c, fl);
                const int side = m->cof(f, 0) == (int32_t)c ? 0 : 1;
                for (int q = 0; q < nq(); q++)
                    for (int v = 0; v < NC; v++)
                        fs(f, q, side, v) = Uc[v] + theta * (fs(f, q, side, v) - Uc[v]);
            }
        }
    }
    // face_reconstruction.cpp:866-1039 (per-cell heap scratch as in :873-874)
    void teno_faces(const double* Ux) {
        const int K = n_dof;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t ci = 0; ci < (int64_t)m->n_cells; ++ci) {
            const uint32_t c = (uint32_t)ci;
            const uint32_t s0 = off_groups[c];
            const int gsz = off_groups[c + 1] - s0;
            const int nfc = m->n_faces_of_cell(c);
            double* sw = new double[gsz * NC];
            double* dw = new double[(size_t)gsz * K * NC];
            double J[4], Ji[4], w0[2], w1[2], w2[2];
            for (int i = 0; i < ND; i++) {
                w0[i] = m->nc(m->node_of_cell(c, 0), i);
                w1[i] = m->nc(m->node_of_cell(c, 1), i);
                w2[i] = m->nc(m->node_of_cell(c, 2), i);
            }
            tri_J(w0, w1, w2, J);
            invert2(J, Ji);
            for (int sl = 0; sl < gsz; ++sl) {
                const uint32_t si = s0 + sl;
                const uint32_t soff = off_st[si];
                const int ssz = off_st[si + 1] - soff;
                const uint64_t roff = off_rm[si];
                if (ssz == 0) {
                    for (int v = 0; v < NC; v++) sw[sl * NC + v] = 0.0;
                    continue;
                }
                double* b = new double[(size_t)ssz * NC];
                for (int nl = 0; nl < ssz; ++nl) {
                    const uint32_t nb = st[soff + nl];
                    for (int v = 0; v < NC; ++v)
                        b[nl * NC + v] = ta[soff + nl] * (Ux[(size_t)nb * NC + v] - Ux[(size_t)c * NC + v]);
                }
                gemm(rm.data() + roff, b, &dw[(size_t)sl * K * NC], K, ssz, ssz, NC, false, false);
                delete[] b;
                double si_[NC];
                double* oia = new double[(size_t)K * NC];
                gemm(OI.data(), &dw[(size_t)sl * K * NC], oia, K, K, K, NC, false, false);
                for (int v = 0; v < NC; ++v) {
                    si_[v] = 0.0;
                    for (int k = 0; k < K; ++k) si_[v] += dw[(size_t)sl * K * NC + k * NC + v] * oia[k * NC + v];
                }
                delete[] oia;
                const double eps = 1.0e-12;
                for (int v = 0; v < NC; v++) sw[sl * NC + v] = 1.0 / std::pow(si_[v] + eps, 6.0);
            }
            for (int v = 0; v < NC; ++v) {
                double sd = 0.0;
                for (int sl = 1; sl < gsz; ++sl) sd += sw[sl * NC + v];
                const double CT1 = 1.0e-7, CT2 = 1.0e-5;
                if (sw[v] / (sd + sw[v]) > CT1) {
                    decisions[(size_t)c * NC + v] = 1;
                    sw[v] = 1.0;
                    for (int sl = 1; sl < gsz; ++sl) sw[sl * NC + v] = 0.0;
                } else {
                    decisions[(size_t)c * NC + v] = 0;
                    if (cfg.teno_variant != ORC_TENO_REFERENCE) sw[v] = 0.0;
                    for (int sl = 1; sl < gsz; ++sl) {
                        if (sw[sl * NC + v] / sd > CT2) sw[sl * NC + v] = (1.0 / K);
                        else if (cfg.teno_variant != ORC_TENO_REFERENCE) sw[sl * NC + v] = 0.0;
                    }
                    sd = 0.0;
                    for (int sl = 1; sl < gsz; ++sl) sd += sw[sl * NC + v];
                    for (int sl = 1; sl < gsz; ++sl) sw[sl * NC + v] /= sd;
                }
                for (int fl = 0; fl < nfc; fl++) {
                    const uint32_t f = m->face_of_cell(c, fl);
                    const int side = m->cof(f, 0) == (int32_t)c ? 0 : 1;
                    const uint32_t a = m->node_of_face(f, 0), bnode = m->node_of_face(f, 1);
                    double x0[2], x1[2];
                    for (int i = 0; i < ND; i++) { x0[i] = m->nc(a, i) - w0[i]; x1[i] = m->nc(bnode, i) - w0[i]; }
                    gemv2(Ji, x0, x0);
                    gemv2(Ji, x1, x1);
                    for (int q = 0; q < nq(); q++) {
                        double& out = fs(f, q, side, v);
                        out = Ux[(size_t)c * NC + v];
                        double xq[2];
                        for (int i = 0; i < ND; i++) xq[i] = (qface.pts[q] + 1.0) * 0.5 * (x1[i] - x0[i]) + x0[i];
                        for (int sl = 0; sl < gsz; sl++) {
                            if (sw[sl * NC + v] == 0.0) continue;
                            const uint32_t soff = off_st[s0 + sl];
                            for (int k = 0; k < K; k++)
                                out += sw[sl * NC + v] * dw[(size_t)sl * K * NC + k * NC + v] *
                                       (cfg.teno_variant == ORC_TENO_CORRECTED
                                            ? (b2(poly[2 * k], poly[2 * k + 1], xq[0], xq[1]) - ipt[k])
                                            : (b2(poly[2 * k], poly[2 * k + 1], xq[0], xq[1]) + (ipt[k] / ta[soff])));
                        }
                    }
                }
            }
            if (cfg.teno_variant == ORC_TENO_CORRECTED) positivity_scale(c, Ux);
            delete[] sw;
            delete[] dw;
        }
    }

    // Canonical mode: each face's area-weighted flux goes to ff, and every
    // cell then adds its faces' contributions in the order of a one-thread run
    // (interior faces in zone order, then each boundary zone in input order),
    // so any thread count gives the one-thread result bit for bit. The default
    // mode adds them with atomics as they come, like the reference.
    bool canonical = false;
    std::vector<double> ff;                 // [n_faces][NC], area * F
    std::vector<uint32_t> cf_off;           // per cell: its entries in cf
    std::vector<int64_t> cf;                // face id, negated plus one for side 0, tagged boundary below
    std::vector<uint8_t> cf_bnd;            // 1 when the entry is a boundary face
    void build_canonical_order() {
        const size_t nc = m->n_cells;
        std::vector<std::vector<std::pair<int64_t, int64_t>>> per(nc);  // (rank, signed face)
        int64_t rank = 0;
        auto add_zone = [&](const std::vector<uint32_t>& faces, bool bnd) {
            for (uint32_t f : faces) {
                const int32_t c0 = m->cof(f, 0), c1 = m->cof(f, 1);
                per[c0].push_back({rank, -(int64_t)f - 1});
                if (!bnd && c1 >= 0) per[c1].push_back({rank, (int64_t)f});
                ++rank;
            }
        };
        add_zone(m->zones[0].faces, false);
        const int64_t first_bnd = rank;
        for (int b = 0; b < cfg.n_bc; ++b) add_zone(m->zones[cfg.bc[b].zone].faces, true);
        cf_off.assign(nc + 1, 0);
        cf.clear();
        cf_bnd.clear();
        for (size_t c = 0; c < nc; ++c) {
            std::sort(per[c].begin(), per[c].end());
            for (auto& e : per[c]) {
                cf.push_back(e.second);
                cf_bnd.push_back(e.first >= first_bnd ? 1 : 0);
            }
            cf_off[c + 1] = (uint32_t)cf.size();
        }
        ff.assign((size_t)m->n_faces * NC, 0.0);
    }

    inline void add_rhs(double* r, uint32_t c, int v, double x) {
        double& dst = r[(size_t)c * NC + v];
        if (nthreads > 1) {
#pragma omp atomic
            dst += x;
        } else {
            dst += x;
        }
    }

    // flux_functor.h:124-178 and the boundary functors
    // (boundary_symmetry.cpp:38-69, boundary_extrapolation.cpp:38-55,
    // boundary_wall_adiabatic.cpp:38-70, boundary_upt.cpp:83-100,
    // boundary_p_out.cpp:57-100). The boundary functors use the solver's
    // Euler instance.
    // Positivity-preserving flux limiter (Hu, Adams & Shu, J. Comput. Phys. 242
    // (2013) 169-180; no reference counterpart): each cell's forward-Euler
    // update is the mean over its S faces of the brackets U - S dt a / V F.
    // With the Rusanov flux of the two cell means (pp_low_flux) every bracket
    // is admissible while S dt a alpha / V <= 1, so each face's flux is
    // blended theta F + (1 - theta) F_lo with the largest theta in [0, 1]
    // keeping both neighbours' brackets at rho, p >= min(1e-13, their
    // theta = 0 values): the density bound in closed form, the pressure (concave
    // along the blend) by 40 bisections. Side 0 first, then side 1. A face
    // whose high-order brackets are both admissible (finite, rho and p at
    // least 1e-13) keeps its flux without the low-order one being formed.
    void pp_low_flux(int bc_type, const orc_bc* bc, const double* nu, const double* U0, const double* U1,
                     double* F) const {
        double Wl[NP], Wr[NP];
        phys.prims(Wl, U0);
        double rr = U0[0];
        if (bc_type < 0) {
            phys.prims(Wr, U1);
            rr = U1[0];
        } else if (bc_type == ORC_BC_SYMMETRY || bc_type == ORC_BC_WALL_ADIABATIC) {
            for (int i = 0; i < NP; i++) Wr[i] = Wl[i];
            const double ul[2] = {Wl[0], Wl[1]};
            const double un = dot2(ul, nu);
            for (int i = 0; i < ND; i++) Wr[i] = ul[i] - 2.0 * un * nu[i];
        } else if (bc_type == ORC_BC_EXTRAPOLATION) {
            phys.prims(Wr, U0);
        } else if (bc_type == ORC_BC_UPT) {
            rr = bc->p / (phys.R * bc->T);
            const double e_bc = phys.cv * bc->T;
            Wr[1] = bc->u[1];
            Wr[0] = bc->u[0];
            Wr[2] = bc->p; Wr[3] = bc->T; Wr[4] = e_bc + bc->p / rr;
        } else if (bc_type == ORC_BC_P_OUT) {
            const double ul[2] = {Wl[0], Wl[1]};
            const double um = norm2_2(ul);
            const double a = phys.sos(Wl[2], U0[0]);
            const double pout = (um < a) ? bc->p : Wl[2];
            const double Tb = Wl[3];
            rr = pout / (phys.R * Tb);
            Wr[0] = ul[0]; Wr[1] = ul[1]; Wr[2] = pout; Wr[3] = Tb; Wr[4] = phys.cv * Tb + pout / rr;
        } else {
            throw std::runtime_error("Unknown boundary type.");
        }
        flux_rusanov(F, nu, U0[0], Wl, Wl[2], phys.gamma, Wl[4], rr, Wr, Wr[2], phys.gamma, Wr[4]);
    }
    double pp_theta(const double* U, double s, const double* F, const double* Flo, double theta) const {
        double B0[NC], B1[NC], w[NC];
        for (int v = 0; v < NC; v++) {
            B0[v] = U[v] + s * Flo[v];
            B1[v] = U[v] + s * F[v];
        }
        const double p0 = pressure_raw(B0);
        if (!(B0[0] > 0.0) || !(p0 > 0.0) || !std::isfinite(B1[0] + B1[1] + B1[2] + B1[3])) return 0.0;
        const double er = std::fmin(1.0e-13, B0[0]), ep = std::fmin(1.0e-13, p0);
        if (B1[0] < er) theta = std::fmin(theta, (B0[0] - er) / (B0[0] - B1[0]));
        for (int v = 0; v < NC; v++) w[v] = B0[v] + theta * (B1[v] - B0[v]);
        if (pressure_raw(w) < ep) {
            double lo = 0.0, hi = theta;
            for (int it = 0; it < 40; it++) {
                const double mid = 0.5 * (lo + hi);
                for (int v = 0; v < NC; v++) w[v] = B0[v] + mid * (B1[v] - B0[v]);
                if (pressure_raw(w) < ep) hi = mid; else lo = mid;
            }
            theta = lo;
        }
        return theta;
    }
    // the high-order bracket U + s F finite with rho and p at least 1e-13
    bool pp_admissible(const double* U, double s, const double* F) const {
        double B[NC];
        for (int v = 0; v < NC; v++) B[v] = U[v] + s * F[v];
        return std::isfinite(B[0] + B[1] + B[2] + B[3]) && B[0] >= 1.0e-13 && pressure_raw(B) >= 1.0e-13;
    }
    void pp_face(uint32_t f, int bc_type, const orc_bc* bc, const double* nu, double* F) const {
        const int32_t c0 = m->cof(f, 0), c1 = m->cof(f, 1);
        const bool bnd = bc_type >= 0;
        const double* U0 = &Ucur[(size_t)c0 * NC];
        const double* U1 = bnd ? U0 : &Ucur[(size_t)c1 * NC];
        const double a = m->face_area[f];
        const double Sd = (double)m->n_faces_of_cell((uint32_t)c0);
        const double V0 = m->cell_volume[c0], V1 = bnd ? V0 : m->cell_volume[c1];
        const double k0 = Sd * pp_dt * a / V0, k1 = Sd * pp_dt * a / V1;
        // F stands when both high-order brackets are admissible
        if (pp_admissible(U0, -k0, F) && (bnd || pp_admissible(U1, k1, F))) return;
        double Flo[NC];
        pp_low_flux(bc_type, bc, nu, U0, U1, Flo);
        double theta = pp_theta(U0, -k0, F, Flo, 1.0);
        if (!bnd) theta = pp_theta(U1, k1, F, Flo, theta);
        if (theta < 1.0)
            for (int v = 0; v < NC; v++) F[v] = theta * F[v] + (1.0 - theta) * Flo[v];
        if (theta < 1.0) ++pp_limited;
    }
    void face_flux(uint32_t f, int bc_type, const orc_bc* bc, double* r) {
        const FluxFn fn = flux_fn(cfg.riemann);
        const double* wq = qface.w.data();
        double nv[2] = {m->face_normals[2 * (size_t)f], m->face_normals[2 * (size_t)f + 1]}, nu[2];
        unit2(nv, nu);
        double F[NC] = {0.0, 0.0, 0.0, 0.0}, Ft[NC];
        for (int q = 0; q < nq(); q++) {
            double Cl[NC], Cr[NC], Wl[NP], Wr[NP];
            for (int v = 0; v < NC; v++) Cl[v] = fs(f, q, 0, v);
            if (bc_type < 0) {  // interior
                for (int v = 0; v < NC; v++) Cr[v] = fs(f, q, 1, v);
                phys.prims(Wl, Cl);
                phys.prims(Wr, Cr);
            } else if (bc_type == ORC_BC_SYMMETRY) {
                for (int v = 0; v < NC; v++) Cr[v] = Cl[v];
                phys.prims(Wl, Cl);
                for (int i = 0; i < NP; i++) Wr[i] = Wl[i];
                const double ul[2] = {Wl[0], Wl[1]};
                const double un = dot2(ul, nu);
                for (int i = 0; i < ND; i++) Wr[i] = ul[i] - 2.0 * un * nu[i];
            } else if (bc_type == ORC_BC_EXTRAPOLATION) {
                // boundary_extrapolation.cpp:38-55: the ghost copies the face
                // state. The CORRECTED variant's ghost is the cell mean (a
                // zero-gradient ghost cell): with the face state the
                // reconstruction's boundary values feed back without upwinding
                // and the four-quadrant problem grows without bound there.
                phys.prims(Wl, Cl);
                if (cfg.recon == ORC_RECON_TENO && cfg.teno_variant == ORC_TENO_CORRECTED) {
                    const int32_t c0 = m->cof(f, 0);
                    for (int v = 0; v < NC; v++) Cr[v] = Ucur[(size_t)c0 * NC + v];
                    phys.prims(Wr, Cr);
                } else {
                    for (int v = 0; v < NC; v++) Cr[v] = Cl[v];
                    for (int i = 0; i < NP; i++) Wr[i] = Wl[i];
                }
            } else if (bc_type == ORC_BC_WALL_ADIABATIC) {
                phys.prims(Wl, Cl);
                Ft[0] = 0.0; Ft[1] = Wl[2] * nu[0]; Ft[2] = Wl[2] * nu[1]; Ft[3] = 0.0;
                for (int v = 0; v < NC; v++) F[v] += wq[q] * Ft[v];
                continue;
            } else if (bc_type == ORC_BC_UPT) {
                phys.prims(Wl, Cl);
                const double rho_bc = bc->p / (phys.R * bc->T);
                const double e_bc = phys.cv * bc->T;
                const double h_bc = e_bc + bc->p / rho_bc;
                Cr[0] = rho_bc;
                Wr[0] = bc->u[0]; Wr[1] = bc->u[1]; Wr[2] = bc->p; Wr[3] = bc->T; Wr[4] = h_bc;
            } else if (bc_type == ORC_BC_P_OUT) {
                phys.prims(Wl, Cl);
                const double rl = Cl[0];
                const double ul[2] = {Wl[0], Wl[1]};
                const double pl = Wl[2], Tl = Wl[3];
                const double um = norm2_2(ul);
                const double a = phys.sos(pl, rl);
                const double pout = (um < a) ? bc->p : pl;
                const double Tb = Tl;
                const double rb = pout / (phys.R * Tb);
                const double eb = phys.cv * Tb;
                const double hb = eb + pout / rb;
                Cr[0] = rb;
                Wr[0] = ul[0]; Wr[1] = ul[1]; Wr[2] = pout; Wr[3] = Tb; Wr[4] = hb;
            } else {
                throw std::runtime_error("Unknown boundary type.");
            }
            fn(Ft, nu, Cl[0], Wl, Wl[2], phys.gamma, Wl[4], Cr[0], Wr, Wr[2], phys.gamma, Wr[4]);
            for (int v = 0; v < NC; v++) F[v] += wq[q] * Ft[v];
        }
        for (int v = 0; v < NC; v++) F[v] *= 0.5;
        if (pp_dt > 0.0) pp_face(f, bc_type, bc, nu, F);
        const double a = m->face_area[f];
        const int32_t c0 = m->cof(f, 0), c1 = m->cof(f, 1);
        if (canonical) {  // -a * F and a * F differ only in sign
            for (int v = 0; v < NC; v++) ff[(size_t)f * NC + v] = a * F[v];
            return;
        }
        for (int v = 0; v < NC; v++) {
            add_rhs(r, c0, v, -a * F[v]);
            if (bc_type < 0 && c1 >= 0) add_rhs(r, c1, v, a * F[v]);
        }
    }

    // solver_rhs.cpp:44-115; parts: bit 0 calc_rhs_interior (:74-108),
    // bit 1 calc_rhs_boundaries (:110-115)
    const double* Ucur = nullptr;  // the state of the residual being evaluated
    double pp_dt = 0.0;            // > 0: the positivity limiter's stage dt (inside take_step)
    mutable std::atomic<int64_t> pp_limited{0};  // faces the limiter changed
    void calc_rhs(const double* Ux, double* r, int parts = 3) {
        Ucur = Ux;
        g_vacuum_guard = cfg.recon == ORC_RECON_TENO && cfg.teno_variant == ORC_TENO_CORRECTED;
        const size_t nc = m->n_cells;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < (int64_t)(nc * NC); ++i) r[i] = 0.0;
        if (cfg.recon == ORC_RECON_FO) fo_faces(Ux); else teno_faces(Ux);
        const auto& zi = m->zones[0].faces;
        if (parts & 1) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int64_t i = 0; i < (int64_t)zi.size(); ++i) face_flux(zi[i], -1, nullptr, r);
        }
        for (int b = 0; b < cfg.n_bc && (parts & 2); ++b) {
            const auto& zf = m->zones[cfg.bc[b].zone].faces;
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int64_t i = 0; i < (int64_t)zf.size(); ++i) face_flux(zf[i], cfg.bc[b].type, &cfg.bc[b], r);
        }
        if (canonical) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int64_t c = 0; c < (int64_t)nc; ++c)
                for (uint32_t e = cf_off[c]; e < cf_off[c + 1]; ++e) {
                    if (!(parts & (cf_bnd[e] ? 2 : 1))) continue;
                    const int64_t sf = cf[e];
                    const size_t f = (size_t)(sf < 0 ? -sf - 1 : sf);
                    for (int v = 0; v < NC; ++v) {
                        const double x = ff[f * NC + v];
                        r[(size_t)c * NC + v] += sf < 0 ? -x : x;
                    }
                }
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t c = 0; c < (int64_t)nc; ++c)
            for (int v = 0; v < NC; ++v) r[(size_t)c * NC + v] /= m->cell_volume[c];
    }

    // solver.cpp:580-742
    double calc_dt(double) {
        const size_t nc = m->n_cells;
        double mx = -1.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(max : mx)
        for (int64_t ci = 0; ci < (int64_t)nc; ++ci) {
            const uint32_t c = (uint32_t)ci;
            double conv = 0.0, acou = 0.0;
            const uint32_t nf = m->n_faces_of_cell(c);
            for (uint32_t l = 0; l < nf; l++) {
                const uint32_t f = m->face_of_cell(c, l);
                int32_t cl = m->cof(f, 0), cr = m->cof(f, 1);
                double nv[2] = {m->face_normals[2 * (size_t)f], m->face_normals[2 * (size_t)f + 1]}, nu[2];
                unit2(nv, nu);
                double s[2];
                if (cr == -1) {
                    s[0] = 2.0 * std::pow(m->cell_volume[cl], 1.0 / ND);
                    s[1] = 2.0 * std::pow(m->cell_volume[cl], 1.0 / ND);
                    cr = cl;
                } else {
                    s[0] = m->cc(cr, 0) - m->cc(cl, 0);
                    s[1] = m->cc(cr, 1) - m->cc(cl, 1);
                }
                const double dxn = std::fabs(dot2(s, nu));
                const double rl = U[(size_t)cl * NC], rr = U[(size_t)cr * NC];
                const double ul[2] = {prim[(size_t)cl * NP], prim[(size_t)cl * NP + 1]};
                const double ur[2] = {prim[(size_t)cr * NP], prim[(size_t)cr * NP + 1]};
                const double pl = prim[(size_t)cl * NP + 2], pr = prim[(size_t)cr * NP + 2];
                const double sl = phys.sos(pl, rl), sr = phys.sos(pr, rr);
                const double sf = 0.5 * (sl + sr);
                const double uf[2] = {0.5 * (ul[0] + ur[0]), 0.5 * (ul[1] + ur[1])};
                const double un = std::fabs(dot2(uf, nu));
                conv += un / dxn;
                const double q = sf / dxn;
                acou += q * q;  // pow(x, 2.0) folds to x * x
            }
            const double g = 3.0 / nf;
            conv *= 1.37 * g;
            acou = 1.37 * std::sqrt(g * acou);
            const double lam = conv + acou;
            mx = std::max(mx, lam);
            cfl_local[c] = lam;
        }
        return mx;
    }

    void update_primitives() {
        const size_t nc = m->n_cells;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t c = 0; c < (int64_t)nc; ++c) phys.prims(&prim[(size_t)c * NP], &U[(size_t)c * NC]);
    }

    // time_integrator.cpp:57-163 then solver.cpp:521-531
    void take_step(double dt) {
        pp_dt = cfg.positivity ? dt : 0.0;  // every stage's forward-Euler substep takes the full dt
        integrate_step(cfg.integrator, dt, nthreads, U, rhs, Utmp,
                       [this](const double* u, double* r) { calc_rhs(u, r); });
        pp_dt = 0.0;
        update_primitives();
        t += dt;
    }

    void set_init(uint32_t c, double rho, double u0, double u1, double p) {
        const double T = p / (rho * phys.R);
        const double e = phys.cv * T;
        const double uu[2] = {u0, u1};
        const double E = e + 0.5 * dot2(uu, uu);
        const double h = e + p / rho;
        double* Uc = &U[(size_t)c * NC];
        Uc[0] = rho; Uc[1] = rho * u0; Uc[2] = rho * u1; Uc[3] = rho * E;
        double* Wc = &prim[(size_t)c * NP];
        Wc[0] = u0; Wc[1] = u1; Wc[2] = p; Wc[3] = T; Wc[4] = h;
    }
};

// ---------------------------------------------------------------- C API
#define ORC_TRY(expr, fail) try { expr; } catch (const std::exception& e) { g_err = e.what(); fail; }

extern "C" {

orc_mesh* orc_mesh_create(int type, uint32_t nx, uint32_t ny, double Lx, double Ly, double jitter,
                          uint64_t seed) {
    auto m = std::make_unique<orc_mesh>();
    m->type = type;
    try {
        if (type == ORC_MESH_CARTESIAN) init_cart(*m, nx, ny, Lx, Ly, jitter, seed);
        else if (type == ORC_MESH_CARTESIAN_TRI) init_cart_tri(*m, nx, ny, Lx, Ly, jitter, seed);
        else if (type == ORC_MESH_WEDGE) init_wedge(*m, nx, ny, Lx, Ly);
        else throw std::runtime_error("MeshType::FILE not implemented.");
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
    return m.release();
}
// A caller's mesh in the reference's array layout (mesh.h:228-253; the
// arrays Mesh::copy_host_to_device uploads), with the geometry computed as
// for the generated meshes (mesh.cpp:167-261).
orc_mesh* orc_mesh_create_arrays(uint32_t n_nodes, const double* node_coords, uint32_t n_cells,
                                 const uint32_t* offsets_nodes_of_cell, const uint32_t* nodes_of_cell,
                                 const uint32_t* faces_of_cell, uint32_t n_faces, const uint32_t* nodes_of_face,
                                 const int32_t* cells_of_face, uint32_t n_zones, const char* const* zone_names,
                                 const uint32_t* zone_offsets, const uint32_t* zone_faces) {
    auto m = std::make_unique<orc_mesh>();
    try {
        if (n_zones < 1 || std::string(zone_names[0]) != "interior")
            throw std::runtime_error("the first zone must be the interior");
        m->type = 0;
        m->n_nodes = n_nodes;
        m->n_cells = n_cells;
        m->n_faces = m->n_faces_real = n_faces;
        m->node_coords.assign(node_coords, node_coords + 2 * (size_t)n_nodes);
        std::vector<std::vector<uint32_t>> noc(n_cells), foc(n_cells), nof(n_faces);
        for (uint32_t c = 0; c < n_cells; ++c)
            for (uint32_t k = offsets_nodes_of_cell[c]; k < offsets_nodes_of_cell[c + 1]; ++k) {
                noc[c].push_back(nodes_of_cell[k]);
                foc[c].push_back(faces_of_cell[k]);
            }
        for (uint32_t f = 0; f < n_faces; ++f) nof[f] = {nodes_of_face[2 * (size_t)f], nodes_of_face[2 * (size_t)f + 1]};
        m->build_csr(noc, foc, nof);
        m->cells_of_face.assign(cells_of_face, cells_of_face + 2 * (size_t)n_faces);
        for (uint32_t z = 0; z < n_zones; ++z)
            m->zones.push_back({zone_names[z], z == 0,
                                std::vector<uint32_t>(zone_faces + zone_offsets[z], zone_faces + zone_offsets[z + 1])});
        m->face_areas();
        m->cell_volumes();
        m->cell_centroids();
        m->face_normals_();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
    return m.release();
}
void orc_mesh_destroy(orc_mesh* m) { delete m; }
void orc_mesh_counts(const orc_mesh* m, int64_t* c) {
    c[0] = m->n_cells; c[1] = m->n_nodes; c[2] = m->n_faces; c[3] = m->n_faces_real;
    c[4] = (int64_t)m->nodes_of_cell.size(); c[5] = (int64_t)m->zones.size();
}
void orc_mesh_arrays(const orc_mesh* m, double* nc, double* cc, double* vol, double* fa, double* fn,
                     uint32_t* noc, uint32_t* onoc, uint32_t* foc, uint32_t* ofoc, uint32_t* nof,
                     uint32_t* onof, int32_t* cof) {
    auto cp = [](auto* dst, const auto& src) { if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0])); };
    cp(nc, m->node_coords); cp(cc, m->cell_coords); cp(vol, m->cell_volume); cp(fa, m->face_area);
    cp(fn, m->face_normals); cp(noc, m->nodes_of_cell); cp(onoc, m->off_noc); cp(foc, m->faces_of_cell);
    cp(ofoc, m->off_foc); cp(nof, m->nodes_of_face); cp(onof, m->off_nof); cp(cof, m->cells_of_face);
}
int64_t orc_mesh_zone(const orc_mesh* m, int z, uint32_t* out, char* name) {
    if (z < 0 || z >= (int)m->zones.size()) return -1;
    const auto& Z = m->zones[z];
    if (out) std::memcpy(out, Z.faces.data(), Z.faces.size() * 4);
    if (name) std::snprintf(name, 32, "%s", Z.name.c_str());
    return (int64_t)Z.faces.size();
}

orc_solver* orc_solver_create(orc_mesh* m, const orc_config* cfg, int n_threads) {
    auto s = std::make_unique<orc_solver>();
    s->m = m;
    s->cfg = *cfg;
    s->nthreads = n_threads < 1 ? 1 : n_threads;
    if (cfg->positivity && cfg->integrator == ORC_TI_RK4) {
        g_err = "the positivity limiter needs forward Euler or SSP-RK3";
        return nullptr;
    }
    ORC_TRY(s->init(), return nullptr);
    return s.release();
}
void orc_solver_destroy(orc_solver* s) { delete s; }
int64_t orc_pp_limited(const orc_solver* s) { return s->pp_limited.load(); }
double orc_pp_theta(const orc_solver* s, const double* U, double sc, const double* F, const double* Flo,
                    double theta) {
    return s->pp_theta(U, sc, F, Flo, theta);
}
int orc_pp_low_flux(const orc_solver* s, int b, const double* nu, const double* U0, const double* U1, double* F) {
    if (b >= s->cfg.n_bc) return 1;
    ORC_TRY(s->pp_low_flux(b < 0 ? -1 : s->cfg.bc[b].type, b < 0 ? nullptr : &s->cfg.bc[b], nu, U0, U1, F),
            return 1);
    return 0;
}
void orc_set_canonical(orc_solver* s, int on) {
    s->canonical = on != 0;
    if (s->canonical) s->build_canonical_order();
    else { std::vector<double>().swap(s->ff); s->cf.clear(); s->cf_off.clear(); s->cf_bnd.clear(); }
}
const char* orc_last_error(void) { return g_err.c_str(); }
void orc_teno_sizes(const orc_solver* s, int64_t* o) {
    o[0] = s->n_dof; o[1] = s->M; o[2] = (int64_t)s->st.size(); o[3] = (int64_t)s->off_st.size();
    o[4] = (int64_t)s->rm.size();
}
void orc_teno_arrays(const orc_solver* s, uint32_t* og, uint32_t* os, uint32_t* st, uint64_t* orm,
                     double* rm, double* ta, double* ipt, double* oi, uint8_t* poly) {
    auto cp = [](auto* dst, const auto& src) { if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0])); };
    cp(og, s->off_groups); cp(os, s->off_st); cp(st, s->st); cp(orm, s->off_rm); cp(rm, s->rm);
    cp(ta, s->ta); cp(ipt, s->ipt); cp(oi, s->OI); cp(poly, s->poly);
}
void orc_init_quadrant(orc_solver* s, double xs, double ys, const double* st) {
    for (uint32_t c = 0; c < s->m->n_cells; ++c) {
        const double x = s->m->cc(c, 0), y = s->m->cc(c, 1);
        const bool r = x >= xs, t = y >= ys;
        const int q = r ? (t ? 0 : 3) : (t ? 1 : 2);  // UR, UL, LL, LR
        s->set_init(c, st[4 * q], st[4 * q + 1], st[4 * q + 2], st[4 * q + 3]);
    }
    s->t = 0.0;
}
void orc_init_constant(orc_solver* s, const double* u, double p, double T) {
    const double rho = p / (s->phys.R * T);
    const double e = s->phys.cv * T;
    const double E = e + 0.5 * dot2(u, u);
    const double h = e + p / rho;
    for (uint32_t c = 0; c < s->m->n_cells; ++c) {
        double* Uc = &s->U[(size_t)c * NC];
        Uc[0] = rho; Uc[1] = rho * u[0]; Uc[2] = rho * u[1]; Uc[3] = rho * E;
        double* Wc = &s->prim[(size_t)c * NP];
        Wc[0] = u[0]; Wc[1] = u[1]; Wc[2] = p; Wc[3] = T; Wc[4] = h;
    }
    s->t = 0.0;
}
void orc_set_state(orc_solver* s, const double* U, const double* prim) {
    std::memcpy(s->U.data(), U, s->U.size() * 8);
    if (prim) std::memcpy(s->prim.data(), prim, s->prim.size() * 8);
    else s->update_primitives();
}
void orc_get_state(const orc_solver* s, double* U, double* prim) {
    if (U) std::memcpy(U, s->U.data(), s->U.size() * 8);
    if (prim) std::memcpy(prim, s->prim.data(), s->prim.size() * 8);
}
int orc_calc_rhs_parts(orc_solver* s, const double* U, int parts, double* rhs) {
    ORC_TRY(s->calc_rhs(U, rhs, parts), return 1);
    return 0;
}

int orc_calc_rhs(orc_solver* s, const double* U, double* rhs) {
    ORC_TRY(s->calc_rhs(U, rhs), return 1);
    return 0;
}
void orc_get_face_values(const orc_solver* s, double* out) {
    std::memcpy(out, s->face.data(), s->face.size() * 8);
}
int orc_calc_dt(orc_solver* s, double cfl, double* dt_out, double* cfl_local) {
    double dt = 0.0;
    ORC_TRY(dt = cfl / s->calc_dt(cfl), return 1);
    for (auto& x : s->cfl_local) x = x * dt;  // KokkosBlas::scal
    if (dt < 0.0) { g_err = "dt negative: " + std::to_string(dt) + "."; return 2; }
    *dt_out = dt;
    if (cfl_local) std::memcpy(cfl_local, s->cfl_local.data(), s->cfl_local.size() * 8);
    return 0;
}
int orc_take_step(orc_solver* s, double dt) {
    ORC_TRY(s->take_step(dt), return 1);
    return 0;
}
int orc_run(orc_solver* s, int n, double cfl, double* dts) {
    for (int i = 0; i < n; ++i) {
        double dt;
        const int rc = orc_calc_dt(s, cfl, &dt, nullptr);
        if (rc) return rc;
        if (dts) dts[i] = dt;
        if (orc_take_step(s, dt)) return 1;
    }
    return 0;
}
double orc_time(const orc_solver* s) { return s->t; }
void orc_teno_decisions(const orc_solver* s, uint8_t* out) {
    std::memcpy(out, s->decisions.data(), s->decisions.size());
}

// RiemannSolver::calc_flux as the reference writes it (no vacuum guard,
// whatever the last residual evaluation used)
void orc_riemann_flux(int type, double* F, const double* n, double rl, const double* ul, double pl,
                      double gl, double hl, double rr, const double* ur, double pr, double gr,
                      double hr) {
    orc_riemann_flux_guarded(type, 0, F, n, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
}
void orc_riemann_flux_guarded(int type, int guard, double* F, const double* n, double rl, const double* ul,
                              double pl, double gl, double hl, double rr, const double* ur, double pr,
                              double gr, double hr) {
    const int saved = g_vacuum_guard;
    g_vacuum_guard = guard != 0;
    flux_fn(type)(F, n, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
    g_vacuum_guard = saved;
}
void orc_physics_constants(double g, double pr, double Tr, double rr, double* o) {
    Euler e;
    e.set(g, pr, Tr, rr, -1e20, 1e20);
    o[0] = e.R; o[1] = e.cp; o[2] = e.cv;
}
void orc_primitives(const double* k, const double* cons, double* prim) {
    Euler e;
    e.gamma = k[0]; e.R = k[1]; e.cp = k[2]; e.cv = k[3]; e.p_min = k[4]; e.p_max = k[5];
    e.prims(prim, cons);
}
void orc_qr_noq(const double* A, double* R, int m, int n) { qr_noq(A, R, m, n); }
void orc_forward_substitution(const double* L, const double* B, double* X, int m, int n, int p,
                              int tL, int tB) { forward_sub(L, B, X, m, n, p, tL, tB); }
void orc_back_substitution(const double* U, const double* B, double* X, int m, int n, int p, int tU,
                           int tB) { back_sub(U, B, X, m, n, p, tU, tB); }
void orc_gemm(const double* A, const double* B, double* C, int m, int n, int p, int q, int tA,
              int tB) { gemm(A, B, C, m, n, p, q, tA, tB); }
void orc_invert2(const double* A, double* Ai) { invert2(A, Ai); }
void orc_invert3(const double* A, double* Ai) { invert3(A, Ai); }
double orc_triangle_area(const double* a, const double* b, const double* c) { return tri_area(a, b, c); }
void orc_triangle_J(const double* a, const double* b, const double* c, double* J) { tri_J(a, b, c, J); }
double orc_basis_1d(int basis, int p, double x) { return basis1(basis, p, x); }
double orc_basis_deriv_1d(int basis, int n, int p, double x) { return basisd(basis, n, p, x); }
void orc_integrator_constant_rhs(int type, double dt, double* U, int n, const double* rhs) {
    // time_integrator_test.cpp:22-28 drives the integrators with an RHS callback
    // that returns a fixed field; the same sequence of BLAS-1 updates is applied.
    std::vector<double> u(U, U + n), k(rhs, rhs + n), tmp;
    auto axpy = [&](double a, const std::vector<double>& x, std::vector<double>& y) { for (int i = 0; i < n; i++) y[i] = y[i] + a * x[i]; };
    if (type == ORC_TI_FE) {
        axpy(dt, k, u);
    } else if (type == ORC_TI_RK4) {
        const double co[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        for (int s = 0; s < 4; s++) axpy(dt * co[s], k, u);
    } else {
        const double co[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
        for (int s = 0; s < 3; s++) axpy(dt * co[s], k, u);
    }
    std::memcpy(U, u.data(), n * 8);
}

// The integrators of the solver on dU/dt = lam * U (n independent entries).
int orc_integrator_linear_rhs(int type, double dt, double lam, double* U, int n) {
    std::vector<double> u(U, U + n), tmp(n);
    std::vector<double> r[4] = {std::vector<double>(n), std::vector<double>(n), std::vector<double>(n),
                                std::vector<double>(n)};
    ORC_TRY(integrate_step(type, dt, 1, u, r, tmp,
                           [&](const double* x, double* y) { for (int i = 0; i < n; i++) y[i] = lam * x[i]; }),
            return 1);
    std::memcpy(U, u.data(), (size_t)n * 8);
    return 0;
}

}  // extern "C"
