// Host numerics shared by the precompute: small dense algebra with the
// reference's operation order, Legendre/monomial bases and quadrature rules.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace mlr {

// common_math.h:145-159
void invert2(const double* A, double* Ai);
// common_math.h:207-213 (in-place safe)
inline void gemv2(const double* A, const double* x, double* y) {
    const double t0 = A[0] * x[0] + A[1] * x[1];
    const double t1 = A[2] * x[0] + A[3] * x[1];
    y[0] = t0;
    y[1] = t1;
}
// common_math.h:534-542
inline void tri_J(const double* v0, const double* v1, const double* v2, double* J) {
    J[0] = v1[0] - v0[0];
    J[1] = v2[0] - v0[0];
    J[2] = v1[1] - v0[1];
    J[3] = v2[1] - v0[1];
}
// common_math.h:510-516
double tri_area(const double* v0, const double* v1, const double* v2);

// R of a Householder QR (common_math.h:355-422) with the reflector applied
// as the reference's dense Q_j product; zero products of the identity part
// are skipped, which leaves every result bit unchanged.
void qr_r(const double* A, double* R, int m, int n, double* work);
// common_math.h:436-457 / 471-493
void forward_sub(const double* L, const double* B, double* X, int m, int n, int p, bool tL, bool tB);
void back_sub(const double* U, const double* B, double* X, int m, int n, int p, bool tU, bool tB);

// basis.h:65-185
double basis_1d(BasisType b, int p, double x);
double basis_deriv_1d(BasisType b, int n, int p, double x);
inline double basis_2d(BasisType b, int px, int py, double x, double y) {
    return basis_1d(b, px, x) * basis_1d(b, py, y);
}

struct QuadRule {
    int dim = 1, order = 1, n = 0;
    std::vector<double> pts, w;
};
QuadRule gauss_legendre(int order);   // quadrature.cpp:23-151
QuadRule dunavant(int order);         // quadrature.cpp:153-273

// Correctly rounded x^6 via a double-double product (Dekker/Knuth exact
// products). It returns the double nearest to x^6, which is what glibc's
// pow(x, 6.0) returns outside its rare misrounded cases.
double pow6_cr(double x);

}  // namespace mlr
