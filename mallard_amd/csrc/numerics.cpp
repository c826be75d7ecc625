#include "numerics.hpp"

#include <cmath>
#include <string>

namespace mlr {

void invert2(const double* A, double* Ai) {
    const double det = A[0] * A[3] - A[1] * A[2];
    if (det == 0.0) fail("Matrix is singular.");
    const double id = 1.0 / det;
    Ai[0] = A[3] * id;
    Ai[1] = -A[1] * id;
    Ai[2] = -A[2] * id;
    Ai[3] = A[0] * id;
}

double tri_area(const double* v0, const double* v1, const double* v2) {
    return 0.5 * std::fabs(v0[0] * (v1[1] - v2[1]) + v1[0] * (v2[1] - v0[1]) + v2[0] * (v0[1] - v1[1]));
}

// work needs m + m*m + m*n doubles.
void qr_r(const double* A, double* R, int m, int n, double* work) {
    for (int i = 0; i < m * n; ++i) R[i] = A[i];
    double* v = work;
    double* Q = work + m;
    double* Rn = Q + m * m;
    for (int j = 0; j < n; j++) {
        double nx = 0.0;
        for (int i = j; i < m; i++) nx += R[i * n + j] * R[i * n + j];
        nx = std::sqrt(nx);
        if (nx < 1.0e-15) continue;
        const double sgn = (R[j * n + j] >= 0.0) ? 1.0 : -1.0;
        const double alpha = -sgn * nx;
        double nu = 0.0;
        const int mj = m - j;
        for (int k = 0; k < mj; k++) {
            v[k] = R[(j + k) * n + j];
            if (k == 0) v[k] -= alpha;
            nu += v[k] * v[k];
        }
        nu = std::sqrt(nu);
        for (int k = 0; k < mj; k++) v[k] /= nu;
        // trailing block of Q_j = I - 2 v v^T (rows/cols >= j)
        for (int a = 0; a < mj; a++)
            for (int b = 0; b < mj; b++) Q[a * mj + b] = (a == b ? 1.0 : 0.0) - 2.0 * v[a] * v[b];
        // rows >= j: the leading zero entries of Q_j add signed zeros to a
        // +0.0 running sum, which leaves it unchanged
        for (int a = 0; a < mj; a++)
            for (int c = 0; c < n; c++) {
                double s = 0.0;
                for (int b = 0; b < mj; b++) s += Q[a * mj + b] * R[(j + b) * n + c];
                Rn[a * n + c] = s;
            }
        // rows < j: identity row, 0.0 + R (only the sign of a zero can change)
        for (int i = 0; i < j; i++)
            for (int c = 0; c < n; c++) R[i * n + c] = R[i * n + c] + 0.0;
        for (int a = 0; a < mj; a++)
            for (int c = 0; c < n; c++) R[(j + a) * n + c] = Rn[a * n + c];
    }
}

void forward_sub(const double* L, const double* B, double* X, int m, int n, int p, bool tL, bool tB) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < p; j++) {
            double s = 0.0;
            for (int k = 0; k < i; k++) s += L[tL ? k * m + i : i * n + k] * X[k * p + j];
            X[i * p + j] = (B[tB ? j * n + i : i * p + j] - s) / L[tL ? i * m + i : i * n + i];
        }
}

void back_sub(const double* U, const double* B, double* X, int m, int n, int p, bool tU, bool tB) {
    for (int i = m - 1; i >= 0; i--)
        for (int j = 0; j < p; j++) {
            double s = 0.0;
            const int end = tU ? n : m;
            for (int k = i + 1; k < end; k++) s += U[tU ? k * m + i : i * n + k] * X[k * p + j];
            X[i * p + j] = (B[tB ? j * n + i : i * p + j] - s) / U[tU ? i * m + i : i * n + i];
        }
}

namespace {
// Standard Legendre polynomials P_0..P_9 in expanded monomial form. Each
// entry: scale, then (coefficient, power) terms in descending power. The
// value is scale * (t_0 +/- t_1 ...), every term evaluated as
// ((|c| * x) * x) ... and combined left to right, as the reference's
// hard-coded expressions evaluate (basis.h:85-94, :117-180).
struct LegTerm { double c; int e; };
struct LegPoly { double scale; int nt; LegTerm t[5]; };
constexpr LegPoly kLegendre[10] = {
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

double legendre(int n, int p, double x) {
    if (p > 9 || n > 9) fail("Legendre basis implemented for p <= 9.");
    if (n > p) return 0.0;
    const LegPoly& P = kLegendre[p];
    double acc = 0.0;
    bool first = true;
    for (int j = 0; j < P.nt; j++) {
        const int e = P.t[j].e;
        if (e < n) continue;
        double term = std::fabs(P.t[j].c);
        for (int k = 0; k < n; k++) term *= (double)(e - k);
        for (int k = 0; k < e - n; k++) term *= x;
        const bool neg = P.t[j].c < 0;
        if (first) { acc = neg ? -term : term; first = false; }
        else acc = neg ? acc - term : acc + term;
    }
    return P.scale * acc;
}
}  // namespace

double basis_1d(BasisType b, int p, double x) {
    if (b == BasisType::Legendre) return legendre(0, p, x);
    return std::pow(x, (double)p);
}

// monomial derivative keeps the reference's (p-k+1) factors (basis.h:72-78)
double basis_deriv_1d(BasisType b, int n, int p, double x) {
    if (b == BasisType::Legendre) return legendre(n, p, x);
    double r = std::pow(x, (double)(p - n));
    for (int k = 0; k < n; ++k) r *= (double)(p - k + 1);
    return r;
}

// Textbook rules with the decimal constants of quadrature.cpp so the parsed
// doubles are the same.
QuadRule gauss_legendre(int order) {
    QuadRule q;
    q.dim = 1;
    q.order = order;
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
    default: fail("Gauss-Legendre quadrature rule of order " + std::to_string(order) + " not implemented.");
    }
    q.n = (int)q.w.size();
    return q;
}

QuadRule dunavant(int order) {
    QuadRule q;
    q.dim = 2;
    q.order = order;
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
    default: fail("Triangle Dunavant quadrature rule of order " + std::to_string(order) + " not implemented.");
    }
    q.n = (int)q.w.size();
    return q;
}

double pow6_cr(double x) {
    // x^2 exactly as hi + lo
    const double h2 = x * x;
    const double l2 = std::fma(x, x, -h2);
    // x^3 = (h2 + l2) * x as a double-double
    const double p3 = h2 * x;
    const double e3 = std::fma(h2, x, -p3) + l2 * x;
    const double h3 = p3 + e3;
    const double l3 = e3 - (h3 - p3);
    // x^6 = (h3 + l3)^2, rounded once
    const double p6 = h3 * h3;
    const double e6 = std::fma(h3, h3, -p6) + 2.0 * h3 * l3;
    return p6 + e6;
}

}  // namespace mlr
