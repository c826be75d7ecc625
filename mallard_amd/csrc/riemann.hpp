// Euler EOS, Riemann solvers and boundary states as device functions, shared
// by the residual kernels (kernels.hip) and the point-wise operators
// (operators.hip). Each follows the reference operation by operation:
//   - primitives:      physics.h:842-867
//   - Riemann fluxes:  riemann_solver.h:205-519 (Roe: textbook, no counterpart)
//   - boundary states: boundary_*.cpp
#pragma once

#include <hip/hip_runtime.h>

#include "ddpow.hpp"
#include "device.hpp"

namespace mlr::dev {
namespace {

// ------------------------------------------------------------ physics
// physics.h:842-867
__device__ inline void prims(const Physics& ph, const double* U, double* W) {
    const double rho = U[0];
    const double u0 = U[1] / rho, u1 = U[2] / rho;
    const double E = U[3] / rho;
    const double e = E - 0.5 * (u0 * u0 + u1 * u1);
    const double p = fmax(ph.p_min, fmin(ph.p_max, (ph.gamma - 1.0) * rho * e));
    const double T = e / ph.cv;
    const double h = e + p / rho;
    W[0] = u0; W[1] = u1; W[2] = p; W[3] = T; W[4] = h;
}

// ------------------------------------------------------------ Riemann
// riemann_solver.h:205-240
__device__ inline void pvrs(const double* Wl, const double* Wr, double& ps) {
    const double rl = Wl[0], ul = Wl[1], pl = Wl[2], gl = Wl[3];
    const double rr = Wr[0], ur = Wr[1], pr = Wr[2], gr = Wr[3];
    const double al = sqrt(gl * pl / rl), ar = sqrt(gr * pr / rr);
    const double ravg = 0.5 * (rl + rr), aavg = 0.5 * (al + ar);
    ps = 0.5 * (pl + pr) + 0.5 * (ul - ur) * ravg * aavg;
}
// riemann_solver.h:242-271 (pressure only; the star densities are unused)
__device__ inline void trrs(const double* Wl, const double* Wr, double& ps, int guard) {
    const double rl = Wl[0], ul = Wl[1], pl = Wl[2], gl = Wl[3];
    const double rr = Wr[0], ur = Wr[1], pr = Wr[2], gr = Wr[3];
    const double al = sqrt(gl * pl / rl), ar = sqrt(gr * pr / rr);
    const double zl = (gl - 1.0) / (2.0 * gl), zr = (gr - 1.0) / (2.0 * gr);
    const double Plr = ddpow::pow_dd((pl / pr), zl);
    const double us = (Plr * ul / al + ur / ar + 2.0 * (1.0 - Plr) / (gl - 1.0)) / (Plr / al + 1.0 / ar);
    double bl = 1.0 + (gl - 1.0) / (2.0 * al) * (ul - us);
    double br = 1.0 + (gr - 1.0) / (2.0 * ar) * (us - ur);
    if (guard) { bl = fmax(bl, 0.0); br = fmax(br, 0.0); }
    ps = 0.5 * (pl * ddpow::pow_dd(bl, (1.0 / zl)) + pr * ddpow::pow_dd(br, (1.0 / zr)));
}
// riemann_solver.h:273-306
__device__ inline void tsrs(const double* Wl, const double* Wr, double& ps) {
    const double rl = Wl[0], ul = Wl[1], pl = Wl[2], gl = Wl[3];
    const double rr = Wr[0], ur = Wr[1], pr = Wr[2], gr = Wr[3];
    const double gml = (gl - 1.0) / (gl + 1.0), gmr = (gr - 1.0) / (gr + 1.0);
    const double Al = 2.0 / (gl + 1.0) / rl, Bl = gml * pl;
    const double Ar = 2.0 / (gr + 1.0) / rr, Br = gmr * pr;
    const double p0 = fmax(0.0, ps);
    const double gl0 = sqrt(Al / (p0 + Bl)), gr0 = sqrt(Ar / (p0 + Br));
    ps = (gl0 * pl + gr0 * pr - (ur - ul)) / (gl0 + gr0);
}
// riemann_solver.h:308-329
__device__ inline double anrs(const double* Wl, const double* Wr, int guard) {
    const double pl = Wl[2], pr = Wr[2];
    const double pmax = fmax(pl, pr), pmin = fmin(pl, pr);
    const double qmax = pmax / pmin, quser = 2.0;
    double ps;
    pvrs(Wl, Wr, ps);
    if ((qmax < quser) && (pmin <= ps) && (ps <= pmax)) {
    } else if (ps < pmin) {
        trrs(Wl, Wr, ps, guard);
    } else {
        tsrs(Wl, Wr, ps);
    }
    return ps;
}

struct LR {
    double uln, urn, al, ar, Ul[4], Ur[4], Fl[4], Fr[4];
};

// riemann_solver.h:338-370
__device__ inline void prelim(LR& s, double n0, double n1, double rl, const double* ul, double pl, double gl,
                              double hl, double rr, const double* ur, double pr, double gr, double hr) {
    s.uln = ul[0] * n0 + ul[1] * n1;
    s.urn = ur[0] * n0 + ur[1] * n1;
    const double ul2 = ul[0] * ul[0] + ul[1] * ul[1], ur2 = ur[0] * ur[0] + ur[1] * ur[1];
    s.al = sqrt(gl * pl / rl);
    s.ar = sqrt(gr * pr / rr);
    const double El = (hl + 0.5 * ul2) * rl - pl, Er = (hr + 0.5 * ur2) * rr - pr;
    s.Ul[0] = rl; s.Ul[1] = rl * ul[0]; s.Ul[2] = rl * ul[1]; s.Ul[3] = El;
    s.Ur[0] = rr; s.Ur[1] = rr * ur[0]; s.Ur[2] = rr * ur[1]; s.Ur[3] = Er;
    s.Fl[0] = rl * s.uln;
    s.Fl[1] = rl * ul[0] * s.uln + pl * n0;
    s.Fl[2] = rl * ul[1] * s.uln + pl * n1;
    s.Fl[3] = (El + pl) * s.uln;
    s.Fr[0] = rr * s.urn;
    s.Fr[1] = rr * ur[0] * s.urn + pr * n0;
    s.Fr[2] = rr * ur[1] * s.urn + pr * n1;
    s.Fr[3] = (Er + pr) * s.urn;
}

__device__ inline void riemann_flux(int type, int guard, double* F, double n0, double n1, double rl,
                                    const double* ul, double pl, double gl, double hl, double rr,
                                    const double* ur, double pr, double gr, double hr) {
    LR s;
    prelim(s, n0, n1, rl, ul, pl, gl, hl, rr, ur, pr, gr, hr);
    if (type == 0) {  // Rusanov, riemann_solver.h:331-375
        const double Smax = fmax(fabs(s.uln) + s.al, fabs(s.urn) + s.ar);
        for (int i = 0; i < 4; i++) F[i] = 0.5 * (s.Fl[i] + s.Fr[i] + Smax * (s.Ul[i] - s.Ur[i]));
        return;
    }
    if (type == 3) {  // Roe with Harten entropy fix (textbook; no reference counterpart)
        const double g = 0.5 * (gl + gr);
        const double Hl = hl + 0.5 * (ul[0] * ul[0] + ul[1] * ul[1]);
        const double Hr = hr + 0.5 * (ur[0] * ur[0] + ur[1] * ur[1]);
        const double sl = sqrt(rl), sr = sqrt(rr), isum = 1.0 / (sl + sr);
        const double u = (sl * ul[0] + sr * ur[0]) * isum;
        const double v = (sl * ul[1] + sr * ur[1]) * isum;
        const double H = (sl * Hl + sr * Hr) * isum;
        const double q2 = u * u + v * v;
        const double a2 = (g - 1.0) * (H - 0.5 * q2);
        const double a = sqrt(fmax(a2, 1e-300));
        const double un = u * n0 + v * n1;
        const double ut = -u * n1 + v * n0;
        const double rho = sl * sr;
        const double dr = rr - rl, dp = pr - pl;
        const double dun = s.urn - s.uln;
        const double dut = (-ur[0] * n1 + ur[1] * n0) - (-ul[0] * n1 + ul[1] * n0);
        const double al1 = (dp - rho * a * dun) / (2.0 * a2);
        const double al2 = dr - dp / a2;
        const double al3 = rho * dut;
        const double al4 = (dp + rho * a * dun) / (2.0 * a2);
        const double d = 0.1 * a;
        auto fix = [&](double lam) {
            const double am = fabs(lam);
            return am < d ? (lam * lam + d * d) / (2.0 * d) : am;
        };
        const double l1 = fix(un - a), l2 = fabs(un), l3 = fabs(un), l4 = fix(un + a);
        const double r1[4] = {1.0, u - a * n0, v - a * n1, H - a * un};
        const double r2[4] = {1.0, u, v, 0.5 * q2};
        const double r3[4] = {0.0, -n1, n0, ut};
        const double r4[4] = {1.0, u + a * n0, v + a * n1, H + a * un};
        for (int i = 0; i < 4; i++)
            F[i] = 0.5 * (s.Fl[i] + s.Fr[i]) -
                   0.5 * (l1 * al1 * r1[i] + l2 * al2 * r2[i] + l3 * al3 * r3[i] + l4 * al4 * r4[i]);
        return;
    }
    const double Wl[4] = {rl, s.uln, pl, gl}, Wr[4] = {rr, s.urn, pr, gr};
    const double ps = anrs(Wl, Wr, guard);
    const double ql = (ps <= pl) ? 1.0 : sqrt(1.0 + (gl + 1.0) / (2.0 * gl) * (ps / pl - 1.0));
    const double qr = (ps <= pr) ? 1.0 : sqrt(1.0 + (gr + 1.0) / (2.0 * gr) * (ps / pr - 1.0));
    const double Sl = s.uln - s.al * ql, Sr = s.urn + s.ar * qr;
    if (type == 1) {  // HLL, riemann_solver.h:377-439
        if (0.0 <= Sl) for (int i = 0; i < 4; i++) F[i] = s.Fl[i];
        else if (Sr <= 0.0) for (int i = 0; i < 4; i++) F[i] = s.Fr[i];
        else for (int i = 0; i < 4; i++)
            F[i] = (Sr * s.Fl[i] - Sl * s.Fr[i] + Sl * Sr * (s.Ur[i] - s.Ul[i])) / (Sr - Sl);
        return;
    }
    // HLLC "variant 2", riemann_solver.h:441-519
    const double Ss = (pr - pl + rl * s.uln * (Sl - s.uln) - rr * s.urn * (Sr - s.urn)) /
                      (rl * (Sl - s.uln) - rr * (Sr - s.urn));
    if (0.0 <= Sl) {
        for (int i = 0; i < 4; i++) F[i] = s.Fl[i];
    } else if (Sr <= 0.0) {
        for (int i = 0; i < 4; i++) F[i] = s.Fr[i];
    } else {
        const double Ds[4] = {0.0, n0, n1, Ss};
        const double Plr = 0.5 * (pl + pr + rl * (Sl - s.uln) * (Ss - s.uln) + rr * (Sr - s.urn) * (Ss - s.urn));
        if (Ss >= 0.0) {
            for (int i = 0; i < 4; i++) F[i] = (Ss * (Sl * s.Ul[i] - s.Fl[i]) + Sl * Plr * Ds[i]) / (Sl - Ss);
        } else {
            for (int i = 0; i < 4; i++) F[i] = (Ss * (Sr * s.Ur[i] - s.Fr[i]) + Sr * Plr * Ds[i]) / (Sr - Ss);
        }
    }
}

// Flux through one face point. Boundary states follow boundary_*.cpp; the
// wall returns the pressure-only flux (boundary_wall_adiabatic.cpp:38-70).
// On a boundary face Cl is the face state and Cr the state an extrapolation
// boundary copies into its ghost (the face state itself in the reference).
__device__ inline void point_flux(const Physics& ph, const BCParam* __restrict__ bct, int riemann, int guard,
                                  bool bnd, int bc, double n0,
                                  double n1, const double* Cl, const double* Cr, double* F) {
    double Wl[5], Wr[5];
    double rr;
    prims(ph, Cl, Wl);
    if (!bnd) {
        prims(ph, Cr, Wr);
        rr = Cr[0];
    } else {
        const BCParam& b = bct[bc];
        if (b.type == 0) {  // symmetry: mirror the velocity (boundary_symmetry.cpp:38-69)
            for (int i = 0; i < 5; i++) Wr[i] = Wl[i];
            const double un = Wl[0] * n0 + Wl[1] * n1;
            Wr[0] = Wl[0] - 2.0 * un * n0;
            Wr[1] = Wl[1] - 2.0 * un * n1;
            rr = Cl[0];
        } else if (b.type == 1) {  // extrapolation: the ghost is the state Cr
            prims(ph, Cr, Wr);
            rr = Cr[0];
        } else if (b.type == 2 || b.type == 5) {  // adiabatic wall (slip; no-slip adds viscous terms)
            F[0] = 0.0; F[1] = Wl[2] * n0; F[2] = Wl[2] * n1; F[3] = 0.0;
            return;
        } else if (b.type == 3) {  // fixed u, p, T (boundary_upt.cpp:83-100)
            rr = b.rho;
            Wr[0] = b.u0; Wr[1] = b.u1; Wr[2] = b.p; Wr[3] = b.T; Wr[4] = b.h;
        } else {  // p_out (boundary_p_out.cpp:57-100)
            const double um = sqrt(Wl[0] * Wl[0] + Wl[1] * Wl[1]);
            const double a = sqrt(ph.gamma * Wl[2] / Cl[0]);
            const double pout = (um < a) ? b.p : Wl[2];
            const double Tb = Wl[3];
            const double rb = pout / (ph.R * Tb);
            const double eb = ph.cv * Tb;
            const double hb = eb + pout / rb;
            rr = rb;
            Wr[0] = Wl[0]; Wr[1] = Wl[1]; Wr[2] = pout; Wr[3] = Tb; Wr[4] = hb;
        }
    }
    riemann_flux(riemann, guard, F, n0, n1, Cl[0], Wl, Wl[2], ph.gamma, Wl[4], rr, Wr, Wr[2], ph.gamma, Wr[4]);
}

// ---------------------------------------------------------------- positivity
// Positivity-preserving flux limiter (Hu, Adams & Shu, J. Comput. Phys. 242
// (2013) 169-180). No reference counterpart. A cell's forward-Euler update
// splits into one bracket per face, U - S dt a / V F, whose mean over the
// cell's S faces is the update. With the first-order Lax-Friedrichs flux of
// the two cell means each bracket is admissible (rho, p > 0) while
// S dt a alpha / V <= 1 (alpha the larger |u.n| + c); the face's flux is
// blended theta F + (1 - theta) F_lo with the largest theta in [0, 1] that
// keeps both neighbours' brackets admissible.

__device__ inline double pp_pressure(const Physics& ph, const double* u) {
    return (ph.gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
}

// The low-order flux: Rusanov (the Lax-Friedrichs flux with the larger local
// wave speed) of the cell means across the face; at a boundary, of the mean
// and the boundary's ghost of it, walls taking the mirror state.
__device__ inline void pp_low_flux(const Physics& ph, const BCParam* __restrict__ bct, bool bnd, int bc, double n0,
                                   double n1, const double* U0, const double* U1, double* F) {
    if (bnd && (bct[bc].type == 2 || bct[bc].type == 5)) {
        double W[5], Wg[5];
        prims(ph, U0, W);
        for (int i = 0; i < 5; i++) Wg[i] = W[i];
        const double un = W[0] * n0 + W[1] * n1;
        Wg[0] = W[0] - 2.0 * un * n0;
        Wg[1] = W[1] - 2.0 * un * n1;
        riemann_flux(0, 0, F, n0, n1, U0[0], W, W[2], ph.gamma, W[4], U0[0], Wg, Wg[2], ph.gamma, Wg[4]);
        return;
    }
    if (bnd) point_flux(ph, bct, 0, 0, true, bc, n0, n1, U0, U0, F);  // two calls: a selected pointer
    else point_flux(ph, bct, 0, 0, false, 0, n0, n1, U0, U1, F);        // would put the means in scratch
}

// The largest theta in [0, theta] for which U + s (F_lo + theta (F - F_lo))
// stays admissible: rho and p at least min(1e-13, their theta = 0 values);
// 0 when even theta = 0 is not admissible or the flux is not finite.
__device__ inline double pp_theta(const Physics& ph, const double* U, double s, const double* F, const double* Flo,
                                  double theta) {
    double B0[4], B1[4], w[4];
    for (int i = 0; i < 4; i++) {
        B0[i] = U[i] + s * Flo[i];
        B1[i] = U[i] + s * F[i];
    }
    const double p0 = pp_pressure(ph, B0);
    if (!(B0[0] > 0.0) || !(p0 > 0.0) || !isfinite(B1[0] + B1[1] + B1[2] + B1[3])) return 0.0;
    const double er = fmin(1.0e-13, B0[0]), ep = fmin(1.0e-13, p0);
    if (B1[0] < er) theta = fmin(theta, (B0[0] - er) / (B0[0] - B1[0]));
    for (int i = 0; i < 4; i++) w[i] = B0[i] + theta * (B1[i] - B0[i]);
    if (pp_pressure(ph, w) < ep) {  // p is concave along the segment: bisect its crossing
        double lo = 0.0, hi = theta;
        for (int it = 0; it < 40; it++) {
            const double mid = 0.5 * (lo + hi);
            for (int i = 0; i < 4; i++) w[i] = B0[i] + mid * (B1[i] - B0[i]);
            if (pp_pressure(ph, w) < ep) hi = mid;
            else lo = mid;
        }
        theta = lo;
    }
    return theta;
}

// Whether U + s F is finite with rho and p at least 1e-13.
__device__ inline bool pp_admissible(const Physics& ph, const double* U, double s, const double* F) {
    double B[4];
    for (int i = 0; i < 4; i++) B[i] = U[i] + s * F[i];
    return isfinite(B[0] + B[1] + B[2] + B[3]) && B[0] >= 1.0e-13 && pp_pressure(ph, B) >= 1.0e-13;
}

// Limits the face flux F (side 0 to side 1, per unit area, unit normal n)
// toward the low-order flux. k0 = S dt a / V0 (side 0 loses a F), k1 =
// S dt a / V1 (side 1 gains it); a boundary face (boundary condition bc) has
// side 0 only. F stands when both high-order brackets are admissible;
// otherwise theta is found side 0 first, then side 1, so both cells of a
// face compute the same value.
__device__ inline void pp_limit(const Physics& ph, const BCParam* __restrict__ bct, bool bnd, int bc, double n0,
                                double n1, double* F, const double* U0, double k0, const double* U1, double k1) {
    if (pp_admissible(ph, U0, -k0, F) && (bnd || pp_admissible(ph, U1, k1, F))) return;
    double Flo[4];
    pp_low_flux(ph, bct, bnd, bc, n0, n1, U0, U1, Flo);
    double theta = pp_theta(ph, U0, -k0, F, Flo, 1.0);
    if (!bnd) theta = pp_theta(ph, U1, k1, F, Flo, theta);
    if (theta < 1.0)
        for (int i = 0; i < 4; i++) F[i] = theta * F[i] + (1.0 - theta) * Flo[i];
}

}  // namespace
}  // namespace mlr::dev
