#include "layout.hpp"

#include "device.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <utility>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace mlr {

// ------------------------------------------------------------ SFC order
namespace {

// Hilbert index of (x, y) on a 2^b x 2^b grid (classic rotate/flip walk).
uint64_t hilbert_d(uint32_t x, uint32_t y, int b) {
    uint64_t d = 0;
    for (uint32_t s = 1u << (b - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
                x &= (s << 1) - 1;
                y &= (s << 1) - 1;
            }
            std::swap(x, y);
        }
    }
    return d;
}

}  // namespace

std::vector<uint32_t> sfc_order(const Mesh& m) {
    const uint32_t n = m.n_cells;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (uint32_t c = 0; c < n; ++c) {
        x0 = std::min(x0, m.cc(c, 0)); x1 = std::max(x1, m.cc(c, 0));
        y0 = std::min(y0, m.cc(c, 1)); y1 = std::max(y1, m.cc(c, 1));
    }
    const int bits = 21;
    const double span = std::max(x1 - x0, y1 - y0);
    const double sc = span > 0 ? ((double)((1u << bits) - 1)) / span : 0.0;
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)n; ++c) {
        const uint32_t qx = (uint32_t)((m.cc(c, 0) - x0) * sc);
        const uint32_t qy = (uint32_t)((m.cc(c, 1) - y0) * sc);
        key[c] = {hilbert_d(qx, qy, bits), (uint32_t)c};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> perm(n);
    for (uint32_t k = 0; k < n; ++k) perm[k] = key[k].second;
    return perm;
}

// ------------------------------------------------------------ stencil search
namespace {

// Open-addressing set of cell ids (the union of rings already built).
struct IdSet {
    std::vector<uint32_t> keys;
    std::vector<uint32_t> used;
    uint32_t mask = 0;
    static constexpr uint32_t kEmpty = 0xFFFFFFFFu;
    IdSet() { keys.assign(4096, kEmpty); mask = 4095; }
    static uint32_t h(uint32_t k) { return (k * 2654435761u) ^ (k >> 16); }
    bool contains(uint32_t k) const {
        for (uint32_t i = h(k) & mask;; i = (i + 1) & mask) {
            if (keys[i] == k) return true;
            if (keys[i] == kEmpty) return false;
        }
    }
    void insert(uint32_t k) {
        if (used.size() * 2 >= keys.size()) grow();
        for (uint32_t i = h(k) & mask;; i = (i + 1) & mask) {
            if (keys[i] == k) return;
            if (keys[i] == kEmpty) { keys[i] = k; used.push_back(i); return; }
        }
    }
    void clear() {
        for (uint32_t i : used) keys[i] = kEmpty;
        used.clear();
    }
    void grow() {
        std::vector<uint32_t> old;
        for (uint32_t i : used) old.push_back(keys[i]);
        keys.assign(keys.size() * 2, kEmpty);
        mask = (uint32_t)keys.size() - 1;
        used.clear();
        for (uint32_t k : old) insert(k);
    }
};

struct RingSearch {
    const Mesh& m;
    IdSet seen;
    std::vector<std::pair<uint32_t, double>> dist;
    explicit RingSearch(const Mesh& mm) : m(mm) {}

    // mesh.cpp:127-165: the cell and its face neighbours, ascending, unique
    int neighbors1(uint32_t c, uint32_t* out) const {
        int n = 0;
        out[n++] = c;
        const uint32_t nf = m.n_faces_of_cell(c);
        for (uint32_t l = 0; l < nf; ++l) {
            const uint32_t f = m.face_of_cell(c, l);
            const int32_t c0 = m.cof(f, 0), c1 = m.cof(f, 1);
            if (c1 == -1) continue;
            out[n++] = c0 == (int32_t)c ? (uint32_t)c1 : (uint32_t)c0;
        }
        std::sort(out, out + n);
        return (int)(std::unique(out, out + n) - out);
    }

    // face_reconstruction.cpp:217-265: cells adjacent to the last ring that no
    // earlier ring holds, in discovery order, then std::sort by squared
    // distance to the target (same element type and comparator, so ties fall
    // the same way).
    void next_ring(const std::vector<uint32_t>& last, uint32_t tgt, std::vector<uint32_t>& ring) {
        ring.clear();
        uint32_t nb[16];
        for (uint32_t c : last) {
            const int k = neighbors1(c, nb);
            for (int i = 0; i < k; ++i)
                if (!seen.contains(nb[i])) {
                    seen.insert(nb[i]);
                    ring.push_back(nb[i]);
                }
        }
        dist.clear();
        const double tx = m.cc(tgt, 0), ty = m.cc(tgt, 1);
        for (uint32_t c : ring) {
            const double dx = m.cc(c, 0) - tx, dy = m.cc(c, 1) - ty;
            double d = 0.0;
            d += dx * dx;
            d += dy * dy;
            dist.push_back(std::make_pair(c, d));
        }
        std::sort(dist.begin(), dist.end(), [](auto& a, auto& b) { return a.second < b.second; });
        for (size_t i = 0; i < ring.size(); ++i) ring[i] = dist[i].first;
    }

    // face_reconstruction.cpp:267-286 (NCB centred stencil)
    void centered(uint32_t c, int M, std::vector<uint32_t>& st) {
        seen.clear();
        st.assign(1, c);
        seen.insert(c);
        std::vector<uint32_t> last{c}, ring;
        while ((int)st.size() < M) {
            next_ring(last, c, ring);
            if (ring.empty()) fail("mesh too small for the requested TENO stencil size");
            for (uint32_t x : ring) {
                if ((int)st.size() == M) break;
                st.push_back(x);
            }
            last.swap(ring);
        }
    }

    // face_reconstruction.cpp:288-404 (Type-4 directional stencils, then the
    // outward fallback for stencils near a boundary)
    void directional(uint32_t c, int M, std::vector<std::vector<uint32_t>>& S) {
        const int ns = (int)m.n_faces_of_cell(c);
        S.assign(ns, {});
        std::vector<std::array<double, 4>> Jinv(ns);
        for (int s = 0; s < ns; ++s) {
            const uint32_t f = m.face_of_cell(c, s);
            const uint32_t a = m.node_of_face(f, 0), b = m.node_of_face(f, 1);
            const double v0[2] = {m.cc(c, 0), m.cc(c, 1)};
            const double v1[2] = {m.nc(a, 0), m.nc(a, 1)};
            const double v2[2] = {m.nc(b, 0), m.nc(b, 1)};
            double J[4];
            tri_J(v0, v1, v2, J);
            invert2(J, Jinv[s].data());
        }
        for (auto& x : S) x.push_back(c);
        seen.clear();
        seen.insert(c);
        std::vector<uint32_t> last{c}, ring;
        std::vector<uint8_t> grew(ns);
        bool done = false;
        while (!done) {
            next_ring(last, c, ring);
            std::fill(grew.begin(), grew.end(), 0);
            for (int s = 0; s < ns; ++s)
                for (uint32_t x : ring) {
                    if ((int)S[s].size() == M) break;
                    const double dx[2] = {m.cc(x, 0) - m.cc(c, 0), m.cc(x, 1) - m.cc(c, 1)};
                    double dt[2];
                    gemv2(Jinv[s].data(), dx, dt);
                    if (!(dt[0] < 0.0) && !(dt[1] < 0.0)) {
                        S[s].push_back(x);
                        grew[s] = 1;
                    }
                }
            done = true;
            for (int s = 0; s < ns; ++s)
                if (m.cof(m.face_of_cell(c, s), 1) != -1 && (int)S[s].size() < M && grew[s]) {
                    done = false;
                    break;
                }
            last.swap(ring);
        }
        for (int s = 0; s < ns; ++s) {
            if (m.cof(m.face_of_cell(c, s), 1) == -1) { S[s].clear(); continue; }
            if ((int)S[s].size() >= M) continue;
            seen.clear();
            for (uint32_t x : S[s]) seen.insert(x);
            std::vector<uint32_t> l2 = S[s];
            while ((int)S[s].size() < M) {
                next_ring(l2, c, ring);
                if (ring.empty()) fail("mesh too small for the requested TENO stencil size");
                for (uint32_t x : ring) {
                    if ((int)S[s].size() == M) break;
                    S[s].push_back(x);
                }
                l2.swap(ring);
            }
        }
    }

    // face_reconstruction.cpp:406-433
    void all(uint32_t c, int M, std::vector<std::vector<uint32_t>>& out) {
        out.clear();
        std::vector<uint32_t> cen;
        centered(c, M, cen);
        out.push_back(cen);
        std::vector<std::vector<uint32_t>> dir;
        directional(c, M, dir);
        for (auto& d : dir) out.push_back(std::move(d));
        for (auto& s : out)
            if ((int)s.size() < M && !s.empty()) fail("Stencil is not full.");
    }
};

int n_dof_of(int order) {
    // face_reconstruction.cpp:170-180
    int n = 1, den = 1;
    for (int i = 1; i <= kDim; ++i) { den *= i; n *= order + i; }
    return n / den;
}

// face_reconstruction.cpp:182-215 (lexicographic index pairs)
std::vector<uint8_t> poly_indices(int order, int K) {
    std::vector<uint8_t> poly(2 * K, 0);
    int idx = 0, ind[2];
    for (int p = 0; p <= order; p++) {
        ind[0] = p; ind[1] = 0;
        poly[2 * idx] = ind[0]; poly[2 * idx + 1] = ind[1]; ++idx;
        while (ind[1] < p) {
            if (ind[0] > 0) { ind[0] -= 1; ind[1] += 1; }
            poly[2 * idx] = ind[0]; poly[2 * idx + 1] = ind[1]; ++idx;
        }
    }
    return poly;
}

struct TenoCtx {
    const Mesh& m;
    TenoSpec ts;
    int K, M;
    std::vector<uint8_t> poly;
    QuadRule qc, qf;
    std::vector<double> ipt;  // from cell 0, stencil 0 (face_reconstruction.cpp:598-602)
    bool have_ipt = false;

    TenoCtx(const Mesh& mm, const TenoSpec& t) : m(mm), ts(t) {
        K = n_dof_of(ts.order);
        M = (uint16_t)(ts.factor * K);
        poly = poly_indices(ts.order, K);
        qc = dunavant(ts.quad_order_cell > 0 ? ts.quad_order_cell : ts.order + 1);
        qf = gauss_legendre(ts.quad_order_face > 0 ? ts.quad_order_face : (ts.order + 1) / 2);
        if (M < 2 || K < 2) fail("TENO needs basis_order >= 1 and more than one cell per stencil");
    }

    void target_frame(uint32_t c, double* w0, double* Ji) const {
        double w1[2], w2[2], J[4];
        for (int i = 0; i < 2; i++) {
            w0[i] = m.nc(m.node_of_cell(c, 0), i);
            w1[i] = m.nc(m.node_of_cell(c, 1), i);
            w2[i] = m.nc(m.node_of_cell(c, 2), i);
        }
        tri_J(w0, w1, w2, J);
        invert2(J, Ji);
    }

    // Basis integrals of one stencil member in the target's frame
    // (face_reconstruction.cpp:525-591): row of K values and the member's
    // transformed area.
    void member_row(uint32_t nb, const double* w0, const double* Ji, double* row, double& at) const {
        double v0[2], v1[2], v2[2], Jn[4];
        for (int i = 0; i < 2; i++) {
            v0[i] = m.nc(m.node_of_cell(nb, 0), i);
            v1[i] = m.nc(m.node_of_cell(nb, 1), i);
            v2[i] = m.nc(m.node_of_cell(nb, 2), i);
        }
        tri_J(v0, v1, v2, Jn);
        double t0[2], t1[2], t2[2];
        for (int i = 0; i < 2; i++) { t0[i] = v0[i] - w0[i]; t1[i] = v1[i] - w0[i]; t2[i] = v2[i] - w0[i]; }
        gemv2(Ji, t0, t0);
        gemv2(Ji, t1, t1);
        gemv2(Ji, t2, t2);
        at = tri_area(t0, t1, t2);
        double qp[2 * 16];
        for (int q = 0; q < qc.n; ++q) {
            double x[2] = {qc.pts[2 * q], qc.pts[2 * q + 1]};
            gemv2(Jn, x, x);
            for (int i = 0; i < 2; i++) x[i] += v0[i];
            for (int i = 0; i < 2; i++) x[i] -= w0[i];
            gemv2(Ji, x, x);
            qp[2 * q] = x[0];
            qp[2 * q + 1] = x[1];
        }
        for (int k = 0; k < K; ++k) {
            double a = 0.0;
            for (int q = 0; q < qc.n; ++q)
                a += qc.w[q] * basis_2d(ts.basis, poly[2 * k], poly[2 * k + 1], qp[2 * q], qp[2 * q + 1]);
            row[k] = a * at;
        }
    }

    // the member area member_row computes for the target itself
    double target_area(uint32_t c, const double* w0, const double* Ji) const {
        double t[3][2];
        for (int l = 0; l < 3; ++l) {
            for (int i = 0; i < 2; i++) t[l][i] = m.nc(m.node_of_cell(c, l), i) - w0[i];
            gemv2(Ji, t[l], t[l]);
        }
        return tri_area(t[0], t[1], t[2]);
    }

    void compute_ipt() {
        RingSearch rs(m);
        std::vector<uint32_t> cen;
        rs.centered(0, M, cen);
        double w0[2], Ji[4];
        target_frame(0, w0, Ji);
        std::vector<double> row(K);
        double at;
        member_row(cen[0], w0, Ji, row.data(), at);
        ipt.resize(K);
        for (int k = 0; k < K; ++k) ipt[k] = row[k] / at;
        have_ipt = true;
    }

    // Pseudo-inverse of one stencil (face_reconstruction.cpp:519-707).
    // aplus: K x M row-major; returns false when the first column of A is not
    // numerically zero (the full-matrix branch was taken).
    bool stencil_matrix(uint32_t c, const std::vector<uint32_t>& st, const double* w0, const double* Ji,
                        double* aplus, double* ta, std::vector<double>& work) const {
        const int n = (int)st.size();
        work.resize((size_t)n * K * 4 + n + n * n + n * K + 64);
        double* A = work.data();
        double* R = A + (size_t)n * K;
        double* Y = R + (size_t)n * K;
        double* B = Y + (size_t)n * K;
        double* qw = B + (size_t)n * K;
        for (int i = 0; i < n; ++i) member_row(st[i], w0, Ji, A + (size_t)i * K, ta[i]);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < K; ++k) A[(size_t)i * K + k] -= ta[i] * ipt[k];
        bool col0 = true;
        for (int i = 0; i < n; ++i)
            if (A[(size_t)i * K] > 1.0e-12) { col0 = false; break; }  // no fabs, as in the reference
        (void)c;
        if (col0) {
            const int mp = n - 1, kp = K - 1;
            for (int i = 1; i < n; ++i)
                for (int k = 1; k < K; ++k) B[(size_t)(i - 1) * kp + (k - 1)] = A[(size_t)i * K + k];
            qr_r(B, R, mp, kp, qw);
            forward_sub(R, B, Y, kp, kp, mp, true, true);
            back_sub(R, Y, B, kp, kp, mp, false, false);
            for (int k = 0; k < K; ++k)
                for (int i = 0; i < n; ++i)
                    aplus[(size_t)k * n + i] = (k == 0 || i == 0) ? 0.0 : B[(size_t)(k - 1) * mp + (i - 1)];
        } else {
            qr_r(A, R, n, K, qw);
            forward_sub(R, A, Y, K, K, n, true, true);
            back_sub(R, Y, A, K, K, n, false, false);
            for (int i = 0; i < n * K; ++i) aplus[i] = A[i];
        }
        return col0;
    }

    // face_reconstruction.cpp:743-788
    std::vector<double> oscillation_indicator() const {
        std::vector<double> OI((size_t)K * K, 0.0);
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) {
                double acc = 0.0;
                for (int k = 1; k < K; k++)
                    for (int q = 0; q < qc.n; ++q) {
                        double di = 1.0, dj = 1.0;
                        for (int d = 0; d < 2; d++)
                            di *= basis_deriv_1d(ts.basis, poly[2 * k + d], poly[2 * i + d], qc.pts[2 * q + d]);
                        for (int d = 0; d < 2; d++)
                            dj *= basis_deriv_1d(ts.basis, poly[2 * k + d], poly[2 * j + d], qc.pts[2 * q + d]);
                        acc += qc.w[q] * di * dj;
                    }
                OI[(size_t)i * K + j] = acc;
            }
        return OI;
    }
};

}  // namespace

void stencils_of_cell(const Mesh& m, int M, uint32_t c, std::vector<std::vector<uint32_t>>& out) {
    RingSearch rs(m);
    rs.all(c, M, out);
}

void teno_matrices_of_cell(const Mesh& m, const TenoSpec& ts, uint32_t c,
                           std::vector<std::vector<double>>& aplus, std::vector<std::vector<double>>& ta,
                           std::vector<double>& ipt_out) {
    TenoCtx ctx(m, ts);
    ctx.compute_ipt();
    ipt_out = ctx.ipt;
    std::vector<std::vector<uint32_t>> S;
    RingSearch rs(m);
    rs.all(c, ctx.M, S);
    double w0[2], Ji[4];
    ctx.target_frame(c, w0, Ji);
    std::vector<double> work;
    aplus.clear();
    ta.clear();
    for (auto& st : S) {
        if (st.empty()) { aplus.emplace_back(); ta.emplace_back(); continue; }
        std::vector<double> a((size_t)ctx.K * st.size()), t(st.size());
        ctx.stencil_matrix(c, st, w0, Ji, a.data(), t.data(), work);
        aplus.push_back(std::move(a));
        ta.push_back(std::move(t));
    }
}

// ------------------------------------------------------------ layout
uint64_t Layout::device_bytes() const {
    auto b = [](const auto& v) { return (uint64_t)v.size() * sizeof(v[0]); };
    return b(volume) + b(nb) + b(info) + b(nx) + b(ny) + b(area) + b(dxn) + b(blk_off) + b(blk_ids) +
           b(st_loc8) + b(st_loc16) + b(st_ta) +
           b(st_A) + b(fbasis);
}

// Per-block member lists and in-list positions (see Layout::blk_off).
void build_tiles(Layout& L) {
    const uint32_t N = L.N, B = Layout::kTileCells;
    const uint32_t nblk = (N + B - 1) / B;
    const int G = L.G, Mp = L.M - 1;
    std::vector<std::vector<int32_t>> lists(nblk);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < (int64_t)nblk; ++b) {
        const uint32_t c0 = (uint32_t)b * B, c1 = std::min(N, c0 + B);
        std::vector<int32_t>& v = lists[b];
        for (uint32_t c = c0; c < c1; ++c) v.push_back((int32_t)c);
        for (int s = 0; s < G; ++s)
            for (int i = 0; i < Mp; ++i)
                for (uint32_t c = c0; c < c1; ++c)
                    if ((L.info[c] >> (20 + s)) & 1u) v.push_back(L.st_ids[L.mem_idx(s, i, c)]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    L.blk_off.assign(nblk + 1, 0);
    uint32_t mx = 0;
    for (uint32_t b = 0; b < nblk; ++b) {
        L.blk_off[b + 1] = L.blk_off[b] + (uint32_t)lists[b].size();
        mx = std::max(mx, (uint32_t)lists[b].size());
    }
    if (mx > Layout::kTileMax16)
        fail("TENO tile of " + std::to_string(mx) + " cells exceeds the LDS limit of " +
             std::to_string(Layout::kTileMax16));
    L.tile_max = mx;
    L.blk_ids.resize(L.blk_off[nblk]);
    const bool small = mx <= Layout::kTileMax8;
    if (small) L.st_loc8.assign((size_t)G * Mp * L.n_pad(), 0);
    else L.st_loc16.assign((size_t)G * Mp * L.n_pad(), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < (int64_t)nblk; ++b) {
        const std::vector<int32_t>& v = lists[b];
        std::copy(v.begin(), v.end(), L.blk_ids.begin() + L.blk_off[b]);
        const uint32_t c0 = (uint32_t)b * B, c1 = std::min(N, c0 + B);
        for (int s = 0; s < G; ++s)
            for (int i = 0; i < Mp; ++i)
                for (uint32_t c = c0; c < c1; ++c) {
                    if (!((L.info[c] >> (20 + s)) & 1u)) continue;
                    const size_t o = L.mem_idx(s, i, c);
                    const uint32_t pos = (uint32_t)(std::lower_bound(v.begin(), v.end(), L.st_ids[o]) - v.begin());
                    if (small) L.st_loc8[o] = (uint8_t)pos;
                    else L.st_loc16[o] = (uint16_t)pos;
                }
    }
}

Layout build_layout(const Mesh& m, Recon recon, const TenoSpec& ts, const std::vector<BoundarySpec>& bcs,
                    int n_threads, bool allow_device) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    (void)n_threads;
    Layout L;
    const uint32_t N = m.n_cells;
    L.N = N;
    L.S = m.max_faces_of_cell();
    if (L.S > 4) fail("cells with more than 4 faces are not supported");
    for (uint32_t c = 0; c < N; ++c)
        if (m.n_faces_of_cell(c) != L.S) fail("mixed cell types are not supported");
    L.perm = sfc_order(m);
    L.iperm.resize(N);
    for (uint32_t j = 0; j < N; ++j) L.iperm[L.perm[j]] = j;

    // accumulation key of each face: interior zone position first, then each
    // boundary condition's zone in input order (solver_rhs.cpp:44-55)
    const uint64_t kNone = ~0ull;
    std::vector<uint64_t> key(m.n_faces, kNone);
    std::vector<uint8_t> fbc(m.n_faces, (uint8_t)kNoBC);
    const int zi = m.zone_index("interior");
    if (zi < 0) fail("mesh has no interior zone");
    for (size_t k = 0; k < m.zones[zi].faces.size(); ++k) key[m.zones[zi].faces[k]] = (uint64_t)k;
    for (size_t b = 0; b < bcs.size(); ++b) {
        const int z = m.zone_index(bcs[b].zone);
        if (z < 0) fail("Boundary name " + bcs[b].zone + " not found in mesh.");
        for (size_t k = 0; k < m.zones[z].faces.size(); ++k) {
            const uint32_t f = m.zones[z].faces[k];
            if (m.cof(f, 1) != -1) fail("boundary zone " + bcs[b].zone + " lists an interior face");
            if (key[f] != kNone) fail("face listed in two boundary conditions (zone " + bcs[b].zone + ")");
            key[f] = ((uint64_t)(b + 1) << 40) | k;
            fbc[f] = (uint8_t)b;
        }
    }
    if (bcs.size() > 14) fail("at most 14 boundary conditions are supported");

    const uint32_t S = L.S;
    L.volume.resize(N);
    L.centroid.resize(2 * (size_t)N);
    L.nb.assign((size_t)S * N, 0);
    L.info.assign((size_t)S * N, 0);
    L.nx.assign((size_t)S * N, 0.0);
    L.ny.assign((size_t)S * N, 0.0);
    L.area.assign((size_t)S * N, 0.0);
    L.dxn.assign((size_t)S * N, 0.0);
    L.dcx.assign((size_t)S * N, 0.0);
    L.dcy.assign((size_t)S * N, 0.0);
    L.mfx.assign((size_t)S * N, 0.0);
    L.mfy.assign((size_t)S * N, 0.0);
    L.face_slot.assign(m.n_faces, ~0u);
    bool bad = false;
#pragma omp parallel for schedule(static)
    for (int64_t jj = 0; jj < (int64_t)N; ++jj) {
        const uint32_t j = (uint32_t)jj, c = L.perm[j];
        L.volume[j] = m.cell_volume[c];
        L.centroid[2 * (size_t)j] = m.cc(c, 0);
        L.centroid[2 * (size_t)j + 1] = m.cc(c, 1);
        uint64_t keys[4];
        for (uint32_t l = 0; l < S; ++l) {
            const uint32_t f = m.face_of_cell(c, l);
            const int32_t c0 = m.cof(f, 0), c1 = m.cof(f, 1);
            const size_t o = (size_t)l * N + j;
            const bool bnd = c1 == -1;
            const uint32_t side = (!bnd && c1 == (int32_t)c) ? 1u : 0u;
            if (!bnd && c0 != (int32_t)c && c1 != (int32_t)c) bad = true;
            uint32_t nbslot = 0, other = c;
            if (!bnd) {
                other = side ? (uint32_t)c0 : (uint32_t)c1;
                for (uint32_t l2 = 0; l2 < m.n_faces_of_cell(other); ++l2)
                    if (m.face_of_cell(other, l2) == f) nbslot = l2;
            } else if (side == 0) {
                L.face_slot[f] = (j << 2) | l;
            }
            if (!bnd && side == 0) L.face_slot[f] = (j << 2) | l;
            L.nb[o] = (int32_t)L.iperm[other];
            uint32_t inf = nbslot | (side << 2);
            if (bnd) inf |= kSlotBoundary | ((uint32_t)fbc[f] << 4);
            L.info[o] = inf;
            // unit normal, as unit<2> (common_math.h:101-106)
            const double n0 = m.face_normals[2 * (size_t)f], n1 = m.face_normals[2 * (size_t)f + 1];
            const double inv = 1 / std::sqrt(n0 * n0 + n1 * n1);
            const double u0 = n0 * inv, u1 = n1 * inv;
            L.nx[o] = u0;
            L.ny[o] = u1;
            L.area[o] = m.face_area[f];
            // spectral-radius length (solver.cpp:662-672; a boundary face
            // takes the reference's own special-case length)
            double s0, s1;
            if (bnd) {
                s0 = 2.0 * std::pow(m.cell_volume[c0], 1.0 / kDim);
                s1 = 2.0 * std::pow(m.cell_volume[c0], 1.0 / kDim);
            } else {
                s0 = m.cc(c1, 0) - m.cc(c0, 0);
                s1 = m.cc(c1, 1) - m.cc(c0, 1);
            }
            L.dxn[o] = std::fabs(s0 * u0 + s1 * u1);
            // viscous face gradients: the vector the directional correction
            // runs along, and the one to the face midpoint
            {
                const uint32_t a = m.node_of_face(f, 0), b = m.node_of_face(f, 1);
                L.mfx[o] = 0.5 * (m.nc(a, 0) + m.nc(b, 0)) - m.cc(c, 0);
                L.mfy[o] = 0.5 * (m.nc(a, 1) + m.nc(b, 1)) - m.cc(c, 1);
            }
            if (bnd) {
                L.dcx[o] = L.mfx[o];
                L.dcy[o] = L.mfy[o];
            } else {
                L.dcx[o] = m.cc(other, 0) - m.cc(c, 0);
                L.dcy[o] = m.cc(other, 1) - m.cc(c, 1);
            }
            keys[l] = key[f];
        }
        // accumulation order: slots by face key, contributing ones first
        uint32_t order[4] = {0, 1, 2, 3};
        std::sort(order, order + S, [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
        uint32_t nact = 0, code = 0;
        for (uint32_t r = 0; r < S; ++r) {
            code |= order[r] << (2 * r);
            if (keys[order[r]] != kNone) ++nact;
        }
        L.info[j] |= (code << 8) | (nact << 16);
    }
    if (bad) fail("inconsistent cells_of_face / faces_of_cell connectivity");

    if (recon == Recon::FO) {
        L.Q = 1;
        L.qface_w = gauss_legendre(1).w;  // face_reconstruction.cpp:41-44
        return L;
    }

    // -------------------------------------------------- TENO precompute
    for (uint32_t c = 0; c < N; ++c)
        if (m.n_nodes_of_cell(c) != 3) fail("TENO has only been implemented for triangular cells.");
    TenoCtx ctx(m, ts);
    ctx.compute_ipt();
    L.teno = true;
    L.tspec = ts;
    L.K = ctx.K;
    L.M = ctx.M;
    L.G = (int)S + 1;
    if (L.G != 4) fail("TENO tables are laid out for triangles (4 stencils per cell)");
    L.Q = (uint32_t)ctx.qf.n;
    L.qface_w = ctx.qf.w;
    L.qface_pts = ctx.qf.pts;
    L.poly = ctx.poly;
    L.ipt = ctx.ipt;
    L.OI = ctx.oscillation_indicator();
    const int K = L.K, M = L.M, G = L.G, Kp = K - 1, Mp = M - 1, Q = (int)L.Q;
    L.basis0 = ts.variant == TenoVariant::Corrected ? basis_2d(ts.basis, 0, 0, 0.0, 0.0) - L.ipt[0] : 1.0;
    const size_t Np = L.n_pad();
    L.device_aplus = allow_device && ts.basis == BasisType::Legendre && dev::aplus_supported(K, M) &&
                     ctx.qc.n <= 8 && K <= 10;
    L.qcell_pts = ctx.qc.pts;
    L.qcell_w = ctx.qc.w;
    L.st_ids.assign((size_t)G * Mp * Np, 0);
    if (!L.device_aplus) {
        L.st_ta.assign((size_t)G * Mp * Np, 0.0);
        L.st_A.assign((size_t)G * Mp * Kp * Np, 0.0);
    }
    L.fbasis.assign((size_t)S * Q * Kp * Np, 0.0);
    int err_code = 0;
    uint32_t err_cell = 0;
#pragma omp parallel
    {
        RingSearch rs(m);
        std::vector<std::vector<uint32_t>> St;
        std::vector<double> work, aplus((size_t)K * M), ta(M);
#pragma omp for schedule(dynamic, 256)
        for (int64_t jj = 0; jj < (int64_t)N; ++jj) {
            const uint32_t j = (uint32_t)jj, c = L.perm[j];
            try {
                rs.all(c, M, St);
            } catch (const std::exception&) {
#pragma omp critical
                { err_code = 1; err_cell = c; }
                continue;
            }
            if (St[0].empty()) {
#pragma omp critical
                { err_code = 2; err_cell = c; }
                continue;
            }
            double w0[2], Ji[4];
            ctx.target_frame(c, w0, Ji);
            double ta0 = NAN;
            uint32_t mask = 0;
            if (L.device_aplus) {
                // stencil ids only; the GPU builds A+ and the areas at upload
                ta0 = ctx.target_area(c, w0, Ji);
                for (int s = 0; s < G; ++s) {
                    if (St[s].empty()) continue;
                    mask |= 1u << s;
                    for (int i = 1; i < M; ++i) L.st_ids[L.mem_idx(s, i - 1, j)] = (int32_t)L.iperm[St[s][i]];
                }
            }
            for (int s = 0; s < G && !L.device_aplus; ++s) {
                if (St[s].empty()) continue;
                mask |= 1u << s;
                const bool col0 = ctx.stencil_matrix(c, St[s], w0, Ji, aplus.data(), ta.data(), work);
                if (!col0) {
#pragma omp critical
                    { err_code = 3; err_cell = c; }
                }
                if (s == 0) ta0 = ta[0];
                else if (ta[0] != ta0) {
#pragma omp critical
                    { err_code = 4; err_cell = c; }
                }
                for (int i = 1; i < M; ++i) {
                    const size_t o = L.mem_idx(s, i - 1, j);
                    L.st_ids[o] = (int32_t)L.iperm[St[s][i]];
                    L.st_ta[o] = ta[i];
                    for (int k = 1; k < K; ++k) L.st_A[L.a_idx(s, i - 1, k - 1, j)] = aplus[(size_t)k * M + i];
                }
            }
            L.info[j] |= mask << 20;
            // face-point basis values in the target frame (face_reconstruction.cpp:985-1031)
            for (uint32_t l = 0; l < S; ++l) {
                const uint32_t f = m.face_of_cell(c, l);
                const uint32_t a = m.node_of_face(f, 0), b = m.node_of_face(f, 1);
                // the face's edge of the reference triangle: the target frame
                // maps the cell's nodes 0, 1, 2 to (0, 0), (1, 0), (0, 1), so
                // the face's midpoint there is the midpoint of edge e = 0
                // (nodes 0-1), 1 (1-2) or 2 (2-0); slot 0's info word holds
                // e of slot l in bits 24 + 2 l
                {
                    int ia = -1, ib = -1;
                    for (int r = 0; r < 3; ++r) {
                        if (m.node_of_cell(c, r) == a) ia = r;
                        if (m.node_of_cell(c, r) == b) ib = r;
                    }
                    if (ia >= 0 && ib >= 0 && ia != ib && l < 3) {
                        const int lo = ia < ib ? ia : ib, hi = ia < ib ? ib : ia;
                        const uint32_t e = (lo == 0 && hi == 1) ? 0u : (lo == 1 && hi == 2) ? 1u : 2u;
                        L.info[j] |= e << (24 + 2 * l);
                        if (ia > ib) L.info[j] |= 1u << dev::face_flip_shift((int)l);
                    }
                }
                double x0[2], x1[2];
                for (int i = 0; i < 2; i++) { x0[i] = m.nc(a, i) - w0[i]; x1[i] = m.nc(b, i) - w0[i]; }
                gemv2(Ji, x0, x0);
                gemv2(Ji, x1, x1);
                for (int q = 0; q < Q; ++q) {
                    double xq[2];
                    for (int i = 0; i < 2; i++) xq[i] = (ctx.qf.pts[q] + 1.0) * 0.5 * (x1[i] - x0[i]) + x0[i];
                    for (int k = 1; k < K; ++k) {
                        const double psi = basis_2d(ts.basis, L.poly[2 * k], L.poly[2 * k + 1], xq[0], xq[1]);
                        const double v = ts.variant == TenoVariant::Corrected ? psi - L.ipt[k] : psi + (L.ipt[k] / ta0);
                        L.fbasis[L.fb_idx(l * Q + q, k - 1, j)] = v;
                    }
                }
            }
        }
    }
    if (err_code == 1) fail("TENO stencil search failed for cell " + std::to_string(err_cell));
    if (err_code == 2) fail("Central stencil is empty for cell " + std::to_string(err_cell) + ".");
    if (err_code == 3)
        fail("TENO stencil of cell " + std::to_string(err_cell) +
             " needs the full-matrix pseudo-inverse branch, which this layout does not store");
    if (err_code == 4) fail("TENO stencils of cell " + std::to_string(err_cell) + " disagree on the target area");
    if (L.teno) build_tiles(L);
    return L;
}

}  // namespace mlr
