// extern "C" boundary (include/mallard_amd.h): converts exceptions and HIP
// failures into status codes plus a per-context error string.
#include "../../include/mallard_amd.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>

#include "ddpow.hpp"
#include "ensemble.hpp"
#include "expr.hpp"
#include "learn.hpp"
#include "solver.hpp"

struct mlr_ctx {
    mlr::Solver solver;
    std::string err;
};

struct mlr_ensemble {
    std::unique_ptr<mlr::Ensemble> e;
};

struct mlr_dataset {
    explicit mlr_dataset(uint64_t cap) : ds(cap) {}
    mlr::Dataset ds;
};

struct mlr_layout {
    mlr::Config cfg;
    mlr::Mesh mesh;
    mlr::Layout L;
};

namespace {

thread_local std::string g_err;

template <class F>
int guarded(mlr_ctx* ctx, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        (ctx ? ctx->err : g_err) = e.what();
        return 1;
    } catch (...) {
        (ctx ? ctx->err : g_err) = "unknown error";
        return 1;
    }
}

mlr::Config config_from_desc(const mlr_numerics_desc* d) {
    using namespace mlr;
    if (!d) fail("numerics descriptor is NULL");
    Config c;
    if (d->recon != MLR_RECON_FO && d->recon != MLR_RECON_TENO) fail("Unknown face reconstruction type.");
    c.recon = d->recon == MLR_RECON_TENO ? Recon::TENO : Recon::FO;
    c.teno.basis = d->basis == MLR_BASIS_LEGENDRE ? BasisType::Legendre : BasisType::Monomial;
    c.teno.order = d->order;
    c.teno.factor = d->stencil_factor > 0 ? d->stencil_factor : 2.0;
    c.teno.quad_order_cell = d->quad_order_cell;
    c.teno.quad_order_face = d->quad_order_face;
    if (d->teno_variant != MLR_TENO_REFERENCE && d->teno_variant != MLR_TENO_CORRECTED) fail("Unknown TENO variant.");
    c.teno.variant = (TenoVariant)d->teno_variant;
    if (d->riemann < 0 || d->riemann > 3) fail("Unknown Riemann solver type.");
    c.riemann = (Riemann)d->riemann;
    if (d->integrator == MLR_TI_FE) c.integrator = Integrator::FE;
    else if (d->integrator == MLR_TI_RK4) c.integrator = Integrator::RK4;
    else if (d->integrator == MLR_TI_SSPRK3) c.integrator = Integrator::SSPRK3;
    else fail("Unknown time integrator type.");
    c.gamma = d->gamma;
    c.p_ref = d->p_ref;
    c.T_ref = d->T_ref;
    c.rho_ref = d->rho_ref;
    c.p_min = d->p_min;
    c.p_max = d->p_max;
    // a nonzero CFL number selects CFL steps (a negative one is then refused
    // with "dt negative" before any step, as solver.cpp:580-590 refuses the
    // negative dt it gives); zero selects the fixed dt
    c.use_cfl = d->cfl != 0.0;
    c.cfl = d->cfl;
    c.dt = d->dt;
    if (d->n_bc < 0 || d->n_bc > 14) fail("n_bc out of range");
    for (int b = 0; b < d->n_bc; ++b) {
        BoundarySpec s;
        s.zone = std::string(d->bc[b].zone, strnlen(d->bc[b].zone, sizeof(d->bc[b].zone)));
        if (d->bc[b].type < 0 || d->bc[b].type > 5) fail("Unknown boundary type.");
        s.type = (BCType)d->bc[b].type;
        s.u[0] = d->bc[b].u[0];
        s.u[1] = d->bc[b].u[1];
        s.p = d->bc[b].p;
        s.T = d->bc[b].T;
        c.bcs.push_back(s);
    }
    c.n_threads = d->n_threads;
    c.viscous = d->viscous != 0;
    if (c.viscous) {
        c.mu_ref = d->mu_ref;
        c.T_mu = d->T_mu_ref > 0.0 ? d->T_mu_ref : c.T_ref;
        c.S_mu = d->sutherland_S;
        c.Pr = d->prandtl > 0.0 ? d->prandtl : 0.72;
        if (!(c.mu_ref >= 0.0)) fail("Invalid viscosity parameters.");
    }
    c.positivity = d->positivity != 0;
    c.n_steps = 0;
    return c;
}

void desc_from_config(const mlr::Config& c, mlr_mesh_gen* mesh, mlr_numerics_desc* d) {
    using namespace mlr;
    if (mesh) {
        std::memset(mesh, 0, sizeof(*mesh));
        mesh->type = (int32_t)c.mesh_type;
        mesh->nx = c.nx;
        mesh->ny = c.ny;
        mesh->Lx = c.Lx;
        mesh->Ly = c.Ly;
        mesh->jitter = c.jitter;
        mesh->seed = c.seed;
    }
    if (d) {
        std::memset(d, 0, sizeof(*d));
        d->recon = c.recon == Recon::TENO ? MLR_RECON_TENO : MLR_RECON_FO;
        d->basis = c.teno.basis == BasisType::Legendre ? MLR_BASIS_LEGENDRE : MLR_BASIS_MONOMIAL;
        d->order = c.teno.order;
        d->stencil_factor = c.teno.factor;
        d->quad_order_cell = c.teno.quad_order_cell;
        d->quad_order_face = c.teno.quad_order_face;
        d->teno_variant = (int32_t)c.teno.variant;
        d->riemann = (int32_t)c.riemann;
        d->integrator = c.integrator == Integrator::FE ? MLR_TI_FE
                                                       : (c.integrator == Integrator::RK4 ? MLR_TI_RK4 : MLR_TI_SSPRK3);
        d->gamma = c.gamma;
        d->p_ref = c.p_ref;
        d->T_ref = c.T_ref;
        d->rho_ref = c.rho_ref;
        d->p_min = c.p_min;
        d->p_max = c.p_max;
        d->cfl = c.use_cfl ? c.cfl : 0.0;
        d->dt = c.dt;
        if (c.bcs.size() > 14) fail("more than 14 boundary conditions");
        d->n_bc = (int32_t)c.bcs.size();
        for (size_t b = 0; b < c.bcs.size(); ++b) {
            std::strncpy(d->bc[b].zone, c.bcs[b].zone.c_str(), sizeof(d->bc[b].zone) - 1);
            d->bc[b].type = (int32_t)c.bcs[b].type;
            d->bc[b].u[0] = c.bcs[b].u[0];
            d->bc[b].u[1] = c.bcs[b].u[1];
            d->bc[b].p = c.bcs[b].p;
            d->bc[b].T = c.bcs[b].T;
        }
        d->n_threads = c.n_threads;
        d->viscous = c.viscous ? 1 : 0;
        d->mu_ref = c.mu_ref;
        d->T_mu_ref = c.T_mu;
        d->sutherland_S = c.S_mu;
        d->prandtl = c.Pr;
        d->positivity = c.positivity ? 1 : 0;
    }
}

}  // namespace

extern "C" {

int mlr_ctx_create_toml(const char* path, mlr_ctx** out) {
    if (!out) return 1;
    *out = nullptr;
    auto ctx = std::make_unique<mlr_ctx>();
    const int rc = guarded(nullptr, [&] { ctx->solver.init(std::string(path ? path : "")); });
    if (rc == 0) *out = ctx.release();
    return rc;
}

int mlr_ctx_create_generated(const mlr_mesh_gen* gen, const mlr_numerics_desc* num, mlr_ctx** out) {
    if (!out) return 1;
    *out = nullptr;
    auto ctx = std::make_unique<mlr_ctx>();
    const int rc = guarded(nullptr, [&] {
        if (!gen) mlr::fail("mesh generator descriptor is NULL");
        mlr::Config c = config_from_desc(num);
        if (gen->type == MLR_MESH_CARTESIAN) c.mesh_type = mlr::MeshType::Cartesian;
        else if (gen->type == MLR_MESH_CARTESIAN_TRI) c.mesh_type = mlr::MeshType::CartesianTri;
        else if (gen->type == MLR_MESH_WEDGE) c.mesh_type = mlr::MeshType::Wedge;
        else mlr::fail("MeshType::FILE not implemented.");
        c.nx = gen->nx;
        c.ny = gen->ny;
        c.Lx = gen->Lx;
        c.Ly = gen->Ly;
        c.jitter = gen->jitter;
        c.seed = gen->seed;
        ctx->solver.init(c);
    });
    if (rc == 0) *out = ctx.release();
    return rc;
}

int mlr_ctx_create(const mlr_mesh_desc* d, const mlr_numerics_desc* num, mlr_ctx** out) {
    if (!out) return 1;
    *out = nullptr;
    auto ctx = std::make_unique<mlr_ctx>();
    const int rc = guarded(nullptr, [&] {
        if (!d) mlr::fail("mesh descriptor is NULL");
        mlr::Config c = config_from_desc(num);
        mlr::Mesh m = mlr::make_from_arrays(d->n_nodes, d->node_coords, d->n_cells, d->offsets_nodes_of_cell,
                                            d->nodes_of_cell, d->faces_of_cell, d->n_faces, d->nodes_of_face,
                                            d->cells_of_face, d->n_zones, d->zone_names, d->zone_offsets,
                                            d->zone_faces);
        ctx->solver.init(c, std::move(m));
    });
    if (rc == 0) *out = ctx.release();
    return rc;
}

int mlr_ctx_destroy(mlr_ctx* ctx) {
    delete ctx;
    return 0;
}

const char* mlr_last_error(const mlr_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int mlr_get_sizes(const mlr_ctx* ctx, int64_t* s) {
    if (!ctx || !s) return 1;
    const mlr::Layout& L = ctx->solver.layout();
    const mlr::Mesh& m = ctx->solver.mesh();
    s[0] = m.n_cells;
    s[1] = m.n_faces;
    s[2] = m.n_faces_real;
    s[3] = L.S;
    s[4] = L.Q;
    s[5] = L.K;
    s[6] = L.M;
    s[7] = L.G;
    s[8] = (int64_t)ctx->solver.device_bytes();
    s[9] = ctx->solver.step();
    s[10] = (int64_t)ctx->solver.config().recon;
    s[11] = (int64_t)L.tspec.variant;
    int64_t n_st = 0;
    if (L.teno)
        for (uint32_t j = 0; j < L.N; ++j) n_st += __builtin_popcount((L.info[j] >> 20) & 0xFu);
    s[12] = n_st;
    s[13] = L.teno ? (int64_t)L.blk_off.back() : 0;       // tile list entries
    s[14] = L.teno ? (ctx->solver.tile_pos_bytes()) : 0;  // bytes per in-tile position
    return 0;
}

static void quadrant_state(mlr::Solver& solver, double xs, double ys, const double* states) {
    const mlr::Mesh& m = solver.mesh();
    const auto& ph_cfg = solver.config();
    const double R = ph_cfg.p_ref / (ph_cfg.T_ref * ph_cfg.rho_ref);
    const double cp = R * ph_cfg.gamma / (ph_cfg.gamma - 1.0);
    const double cv = cp / ph_cfg.gamma;
    std::vector<double> U(4 * (size_t)m.n_cells), W(5 * (size_t)m.n_cells);
    for (uint32_t c2 = 0; c2 < m.n_cells; ++c2) {
        const bool r = m.cc(c2, 0) >= xs, t = m.cc(c2, 1) >= ys;
        const int q = r ? (t ? 0 : 3) : (t ? 1 : 2);
        const double rho = states[4 * q], u0 = states[4 * q + 1], u1 = states[4 * q + 2], p = states[4 * q + 3];
        const double T = p / (rho * R);
        const double e = cv * T;
        const double E = e + 0.5 * (u0 * u0 + u1 * u1);
        const double h = e + p / rho;
        double* Uc = &U[4 * (size_t)c2];
        Uc[0] = rho; Uc[1] = rho * u0; Uc[2] = rho * u1; Uc[3] = rho * E;
        double* Wc = &W[5 * (size_t)c2];
        Wc[0] = u0; Wc[1] = u1; Wc[2] = p; Wc[3] = T; Wc[4] = h;
    }
    solver.set_state(U.data(), W.data());
}

int mlr_init_quadrant(mlr_ctx* ctx, double xs, double ys, const double* states) {
    if (!ctx || !states) return 1;
    return guarded(ctx, [&] { quadrant_state(ctx->solver, xs, ys, states); });
}

int mlr_init_constant(mlr_ctx* ctx, const double* u, double p, double T) {
    if (!ctx || !u) return 1;
    return guarded(ctx, [&] {
        const mlr::Mesh& m = ctx->solver.mesh();
        const auto& cf = ctx->solver.config();
        const double R = cf.p_ref / (cf.T_ref * cf.rho_ref);
        const double cp = R * cf.gamma / (cf.gamma - 1.0);
        const double cv = cp / cf.gamma;
        const double rho = p / (R * T);
        const double e = cv * T;
        const double E = e + 0.5 * (u[0] * u[0] + u[1] * u[1]);
        const double h = e + p / rho;
        std::vector<double> U(4 * (size_t)m.n_cells), W(5 * (size_t)m.n_cells);
        for (uint32_t c2 = 0; c2 < m.n_cells; ++c2) {
            double* Uc = &U[4 * (size_t)c2];
            Uc[0] = rho; Uc[1] = rho * u[0]; Uc[2] = rho * u[1]; Uc[3] = rho * E;
            double* Wc = &W[5 * (size_t)c2];
            Wc[0] = u[0]; Wc[1] = u[1]; Wc[2] = p; Wc[3] = T; Wc[4] = h;
        }
        ctx->solver.set_state(U.data(), W.data());
    });
}

int mlr_set_state(mlr_ctx* ctx, const double* U, const double* prim) {
    if (!ctx || !U) return 1;
    return guarded(ctx, [&] { ctx->solver.set_state(U, prim); });
}

int mlr_get_state(mlr_ctx* ctx, double* U, double* prim) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.get_state(U, prim); });
}

int mlr_calc_rhs(mlr_ctx* ctx, double* rhs) {
    if (!ctx || !rhs) return 1;
    return guarded(ctx, [&] { ctx->solver.calc_rhs(rhs); });
}

int mlr_compute_dt(mlr_ctx* ctx, double* dt_out) {
    if (!ctx) return 1;
    return guarded(ctx, [&] {
        const double dt = ctx->solver.calc_dt();
        if (dt_out) *dt_out = dt;
    });
}

int mlr_step(mlr_ctx* ctx) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.take_step(); });
}

int mlr_run_steps(mlr_ctx* ctx, int64_t n, double* dts) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.run_steps(n, dts); });
}

namespace {
void tune_out(const mlr::dev::Tune& t, int32_t* out) {
    out[0] = t.teno_occ;
    out[1] = (int32_t)t.flux_block;
    out[2] = (int32_t)t.spec_block;
    out[3] = t.teno_occ_reuse;
    out[4] = t.teno_face_basis;
}
mlr::dev::Tune tune_in(const int32_t* in) {
    mlr::dev::Tune t;
    t.teno_occ = in[0];
    t.flux_block = (uint32_t)in[1];
    t.spec_block = (uint32_t)in[2];
    t.teno_occ_reuse = in[3];
    t.teno_face_basis = in[4];
    return t;
}
}  // namespace

int mlr_run_steps_profiled(mlr_ctx* ctx, int64_t n, double* stats) {
    if (!ctx || !stats) return 1;
    return guarded(ctx, [&] { ctx->solver.run_steps_profiled(n, stats); });
}
int mlr_autotune(mlr_ctx* ctx, int32_t* tune) {
    if (!ctx || !tune) return 1;
    return guarded(ctx, [&] {
        ctx->solver.autotune();
        tune_out(ctx->solver.tune(), tune);
    });
}
int mlr_set_tune(mlr_ctx* ctx, const int32_t* tune) {
    if (!ctx || !tune) return 1;
    return guarded(ctx, [&] { ctx->solver.set_tune(tune_in(tune)); });
}
int mlr_get_tune(mlr_ctx* ctx, int32_t* tune) {
    if (!ctx || !tune) return 1;
    return guarded(ctx, [&] { tune_out(ctx->solver.tune(), tune); });
}

int mlr_run(mlr_ctx* ctx) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.run(); });
}

int mlr_set_cfl(mlr_ctx* ctx, double cfl) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.set_cfl(cfl); });
}

int mlr_set_dt_history(mlr_ctx* ctx, int64_t steps) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.set_dt_history(steps); });
}

int mlr_get_time(const mlr_ctx* ctx, double* t, int64_t* step) {
    if (!ctx) return 1;
    if (t) *t = ctx->solver.time();
    if (step) *step = ctx->solver.step();
    return 0;
}

int mlr_get_face_values(mlr_ctx* ctx, double* out) {
    if (!ctx || !out) return 1;
    return guarded(ctx, [&] { ctx->solver.get_face_values(out); });
}

int mlr_calc_face_values(mlr_ctx* ctx, const double* U, double* face_solution) {
    if (!ctx || !U || !face_solution) return 1;
    return guarded(ctx, [&] { ctx->solver.calc_face_values_of(U, face_solution); });
}

int mlr_calc_rhs_of(mlr_ctx* ctx, const double* U, int parts, double* rhs) {
    if (!ctx || !U || !rhs) return 1;
    return guarded(ctx, [&] { ctx->solver.calc_rhs_of(U, parts, rhs); });
}

int mlr_calc_rhs_device(mlr_ctx* ctx, const double* U_dev, int parts, double* rhs_dev) {
    if (!ctx || !U_dev || !rhs_dev) return 1;
    return guarded(ctx, [&] { ctx->solver.calc_rhs_device(U_dev, parts, rhs_dev); });
}

int mlr_riemann_flux(int type, int guard, int64_t n, const double* n_unit, const double* left, const double* right,
                     double* flux) {
    return guarded(nullptr, [&] { mlr::dev::riemann_flux_host(type, guard, n, n_unit, left, right, flux); });
}

int mlr_riemann_flux_noalloc(int type, int guard, int64_t n, const double* n_unit, const double* left,
                             const double* right, double* flux) {
    return guarded(nullptr, [&] { mlr::dev::riemann_flux_noalloc(type, guard, n, n_unit, left, right, flux); });
}

int64_t mlr_riemann_flux_noalloc_capacity(void) { return mlr::dev::riemann_flux_noalloc_capacity(); }

// time_integrator.cpp:57-163: the reference's sequence of rhs evaluations,
// deep copies and BLAS-1 updates, each update one kernel on the stream
int mlr_integrator_take_step(mlr_ctx* ctx, int type, double dt, int64_t n, double* const* solution_vec,
                             double* const* rhs_vec, mlr_rhs_fn rhs, void* user) {
    return guarded(ctx, [&] {
        using mlr::fail;
        if (type != MLR_TI_FE && type != MLR_TI_RK4 && type != MLR_TI_SSPRK3) fail("Unknown time integrator type.");
        if (n < 0) fail("take_step: negative vector length");
        if (!solution_vec || !rhs_vec || !rhs) fail("take_step: NULL vector list or rhs function");
        const int n_sol = type == MLR_TI_FE ? 1 : 2, n_rhs = type == MLR_TI_FE ? 1 : (type == MLR_TI_RK4 ? 4 : 3);
        for (int i = 0; i < n_sol; ++i)
            if (!solution_vec[i]) fail("take_step: NULL solution vector");
        for (int i = 0; i < n_rhs; ++i)
            if (!rhs_vec[i]) fail("take_step: NULL rhs vector");
        hipStream_t st = ctx ? ctx->solver.stream() : nullptr;
        auto chk = [](hipError_t e) { if (e != hipSuccess) fail(hipGetErrorString(e)); };
        auto calc = [&](const double* U, double* r) {
            if (rhs(U, r, user) != 0) fail("take_step: the rhs function failed");
        };
        auto copy = [&](const double* x, double* y) { chk(mlr::dev::vec_update(0, n, 0.0, x, 0.0, y, st)); };
        auto axpy = [&](double a, const double* x, double* y) { chk(mlr::dev::vec_update(1, n, a, x, 0.0, y, st)); };
        auto axpby = [&](double a, const double* x, double b, double* y) {
            chk(mlr::dev::vec_update(2, n, a, x, b, y, st));
        };
        double* U = solution_vec[0];
        if (type == MLR_TI_FE) {
            calc(U, rhs_vec[0]);
            axpy(dt, rhs_vec[0], U);
        } else if (type == MLR_TI_RK4) {
            double* Ut = solution_vec[1];
            const double c[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
            calc(U, rhs_vec[0]);
            copy(U, Ut);
            axpy(dt / 2.0, rhs_vec[0], Ut);
            calc(Ut, rhs_vec[1]);
            copy(U, Ut);
            axpy(dt / 2.0, rhs_vec[1], Ut);
            calc(Ut, rhs_vec[2]);
            copy(U, Ut);
            axpy(dt, rhs_vec[2], Ut);
            calc(Ut, rhs_vec[3]);
            for (int k = 0; k < 4; ++k) axpy(dt * c[k], rhs_vec[k], U);
        } else {
            double* Ut = solution_vec[1];
            const double c[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
            calc(U, rhs_vec[0]);
            copy(U, Ut);
            axpy(dt, rhs_vec[0], Ut);
            calc(Ut, rhs_vec[1]);
            axpby(3.0 / 4.0, U, 1.0 / 4.0, Ut);
            axpy(dt / 4.0, rhs_vec[1], Ut);
            calc(Ut, rhs_vec[2]);
            for (int k = 0; k < 3; ++k) axpy(dt * c[k], rhs_vec[k], U);
        }
        chk(hipStreamSynchronize(st));
    });
}

int mlr_compact_tables(mlr_ctx* ctx) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.compact_tables(); });
}

int mlr_get_decisions(mlr_ctx* ctx, uint8_t* out) {
    if (!ctx || !out) return 1;
    return guarded(ctx, [&] { ctx->solver.get_decisions(out); });
}

int mlr_get_cfl_local(mlr_ctx* ctx, double* out) {
    if (!ctx || !out) return 1;
    return guarded(ctx, [&] { ctx->solver.get_cfl_local(out); });
}

int mlr_get_permutation(const mlr_ctx* ctx, uint32_t* perm) {
    if (!ctx || !perm) return 1;
    const auto& p = ctx->solver.layout().perm;
    std::memcpy(perm, p.data(), p.size() * sizeof(uint32_t));
    return 0;
}

int mlr_get_stencils(const mlr_ctx* ctx, uint32_t cell, uint32_t* ids, uint32_t* sizes) {
    if (!ctx) return 1;
    return guarded(const_cast<mlr_ctx*>(ctx), [&] {
        const mlr::Layout& L = ctx->solver.layout();
        if (!L.teno) mlr::fail("stencils exist only for TENO");
        if (cell >= ctx->solver.mesh().n_cells) mlr::fail("cell out of range");
        std::vector<std::vector<uint32_t>> S;
        mlr::stencils_of_cell(ctx->solver.mesh(), L.M, cell, S);
        for (int s = 0; s < L.G; ++s) {
            sizes[s] = (uint32_t)S[s].size();
            for (int i = 0; i < L.M; ++i) ids[s * L.M + i] = i < (int)S[s].size() ? S[s][i] : 0xFFFFFFFFu;
        }
    });
}

int mlr_get_teno_matrices(const mlr_ctx* ctx, uint32_t cell, double* aplus, double* ta, double* ipt,
                          double* oi) {
    if (!ctx) return 1;
    return guarded(const_cast<mlr_ctx*>(ctx), [&] {
        const mlr::Layout& L = ctx->solver.layout();
        if (!L.teno) mlr::fail("matrices exist only for TENO");
        if (cell >= ctx->solver.mesh().n_cells) mlr::fail("cell out of range");
        std::vector<std::vector<double>> A, T;
        std::vector<double> ip;
        mlr::teno_matrices_of_cell(ctx->solver.mesh(), L.tspec, cell, A, T, ip);
        const int K = L.K, M = L.M;
        for (int s = 0; s < L.G; ++s)
            for (int i = 0; i < K * M; ++i) aplus[(size_t)s * K * M + i] = A[s].empty() ? 0.0 : A[s][i];
        for (int s = 0; s < L.G; ++s)
            for (int i = 0; i < M; ++i) ta[s * M + i] = T[s].empty() ? 0.0 : T[s][i];
        if (ipt) for (int k = 0; k < K; ++k) ipt[k] = ip[k];
        if (oi) std::memcpy(oi, L.OI.data(), L.OI.size() * sizeof(double));
    });
}

int mlr_get_device_matrices(const mlr_ctx* ctx, uint32_t cell, double* aplus, double* ta) {
    if (!ctx || !aplus || !ta) return 1;
    return guarded(const_cast<mlr_ctx*>(ctx), [&] { ctx->solver.device_matrices(cell, aplus, ta); });
}

int mlr_get_device_stencils(const mlr_ctx* ctx, uint32_t cell, uint32_t* ids) {
    if (!ctx) return 1;
    return guarded(const_cast<mlr_ctx*>(ctx), [&] { ctx->solver.device_stencils(cell, ids); });
}

int mlr_layout_create(const mlr_mesh_gen* gen, const mlr_numerics_desc* num, mlr_layout** out) {
    if (!out) return 1;
    *out = nullptr;
    auto lay = std::make_unique<mlr_layout>();
    const int rc = guarded(nullptr, [&] {
        if (!gen) mlr::fail("mesh generator descriptor is NULL");
        lay->cfg = config_from_desc(num);
        if (gen->type == MLR_MESH_CARTESIAN) lay->mesh = mlr::make_cartesian(gen->nx, gen->ny, gen->Lx, gen->Ly, gen->jitter, gen->seed);
        else if (gen->type == MLR_MESH_CARTESIAN_TRI) lay->mesh = mlr::make_cartesian_tri(gen->nx, gen->ny, gen->Lx, gen->Ly, gen->jitter, gen->seed);
        else if (gen->type == MLR_MESH_WEDGE) lay->mesh = mlr::make_wedge(gen->nx, gen->ny, gen->Lx, gen->Ly);
        else mlr::fail("MeshType::FILE not implemented.");
        lay->L = mlr::build_layout(lay->mesh, lay->cfg.recon, lay->cfg.teno, lay->cfg.bcs, lay->cfg.n_threads);
    });
    if (rc == 0) *out = lay.release();
    return rc;
}

int mlr_layout_destroy(mlr_layout* lay) {
    delete lay;
    return 0;
}

int mlr_layout_sizes(const mlr_layout* lay, int64_t* s) {
    if (!lay || !s) return 1;
    const mlr::Layout& L = lay->L;
    s[0] = lay->mesh.n_cells;
    s[1] = lay->mesh.n_faces;
    s[2] = lay->mesh.n_faces_real;
    s[3] = L.S;
    s[4] = L.Q;
    s[5] = L.K;
    s[6] = L.M;
    s[7] = L.G;
    s[8] = (int64_t)L.device_bytes();
    s[9] = lay->mesh.n_nodes;
    s[10] = (int64_t)lay->mesh.nodes_of_cell.size();
    s[11] = (int64_t)lay->mesh.zones.size();
    s[12] = L.tile_max;
    s[13] = L.st_loc16.empty() ? 1 : 2;
    return 0;
}

int mlr_layout_permutation(const mlr_layout* lay, uint32_t* perm) {
    if (!lay || !perm) return 1;
    std::memcpy(perm, lay->L.perm.data(), lay->L.perm.size() * 4);
    return 0;
}

int mlr_layout_mesh_arrays(const mlr_layout* lay, double* nc, double* cc, double* vol, double* fa, double* fn,
                           uint32_t* noc, uint32_t* foc, uint32_t* nof, int32_t* cof) {
    if (!lay) return 1;
    const mlr::Mesh& m = lay->mesh;
    auto cp = [](auto* dst, const auto& src) { if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0])); };
    cp(nc, m.node_coords); cp(cc, m.cell_coords); cp(vol, m.cell_volume); cp(fa, m.face_area);
    cp(fn, m.face_normals); cp(noc, m.nodes_of_cell); cp(foc, m.faces_of_cell); cp(nof, m.nodes_of_face);
    cp(cof, m.cells_of_face);
    return 0;
}

int mlr_layout_cell_stencils(const mlr_layout* lay, uint32_t cell, uint32_t* ids) {
    if (!lay || !ids) return 1;
    return guarded(nullptr, [&] {
        const mlr::Layout& L = lay->L;
        if (!L.teno) mlr::fail("stencils exist only for TENO");
        if (cell >= L.N) mlr::fail("cell out of range");
        const uint32_t j = L.iperm[cell];
        const int Mp = L.M - 1;
        // read back through the uploaded form: the block's tile list at the
        // member's stored position
        const uint32_t b = j / mlr::Layout::kTileCells;
        for (int s = 0; s < L.G; ++s) {
            const bool on = (L.info[j] >> (20 + s)) & 1u;
            ids[(size_t)s * L.M] = on ? cell : 0xFFFFFFFFu;
            for (int i = 0; i < Mp; ++i) {
                const size_t o = L.mem_idx(s, i, j);
                const uint32_t pos = L.st_loc8.empty() ? L.st_loc16[o] : L.st_loc8[o];
                if (on && L.blk_off[b] + pos >= L.blk_off[b + 1]) mlr::fail("tile position out of range");
                ids[(size_t)s * L.M + 1 + i] = on ? L.perm[L.blk_ids[L.blk_off[b] + pos]] : 0xFFFFFFFFu;
            }
        }
    });
}

int mlr_layout_cell_matrices(const mlr_layout* lay, uint32_t cell, double* aplus, double* ta, double* fbasis) {
    if (!lay) return 1;
    return guarded(nullptr, [&] {
        const mlr::Layout& L = lay->L;
        if (!L.teno) mlr::fail("matrices exist only for TENO");
        if (cell >= L.N) mlr::fail("cell out of range");
        const uint32_t j = L.iperm[cell];
        const int K = L.K, M = L.M, Kp = K - 1;
        // the target's transformed area is identical in all stencils; the
        // reference variant folds it into fbasis, so it is recovered from
        // the cell geometry here
        double w0[2], w1[2], w2[2], J[4], Ji[4];
        const mlr::Mesh& m = lay->mesh;
        for (int i = 0; i < 2; ++i) {
            w0[i] = m.nc(m.node_of_cell(cell, 0), i);
            w1[i] = m.nc(m.node_of_cell(cell, 1), i);
            w2[i] = m.nc(m.node_of_cell(cell, 2), i);
        }
        mlr::tri_J(w0, w1, w2, J);
        mlr::invert2(J, Ji);
        double t0[2] = {0.0, 0.0}, t1[2], t2[2];
        for (int i = 0; i < 2; ++i) { t0[i] = w0[i] - w0[i]; t1[i] = w1[i] - w0[i]; t2[i] = w2[i] - w0[i]; }
        mlr::gemv2(Ji, t0, t0);
        mlr::gemv2(Ji, t1, t1);
        mlr::gemv2(Ji, t2, t2);
        const double ta0 = mlr::tri_area(t0, t1, t2);
        for (int s = 0; s < L.G; ++s) {
            const bool on = (L.info[j] >> (20 + s)) & 1u;
            for (int k = 0; k < K; ++k)
                for (int i = 0; i < M; ++i) {
                    double v = 0.0;
                    if (on && k > 0 && i > 0) v = L.st_A[L.a_idx(s, i - 1, k - 1, j)];
                    if (aplus) aplus[((size_t)s * K + k) * M + i] = v;
                }
            if (ta)
                for (int i = 0; i < M; ++i)
                    ta[(size_t)s * M + i] = !on ? 0.0 : (i == 0 ? ta0 : L.st_ta[L.mem_idx(s, i - 1, j)]);
        }
        if (fbasis)
            for (uint32_t l = 0; l < L.S; ++l)
                for (uint32_t q = 0; q < L.Q; ++q)
                    for (int k = 0; k < Kp; ++k)
                        fbasis[((size_t)l * L.Q + q) * Kp + k] = L.fbasis[L.fb_idx(l * L.Q + q, k, j)];
    });
}

int mlr_layout_teno_tables(const mlr_layout* lay, double* ipt, double* oi) {
    if (!lay) return 1;
    if (ipt) std::memcpy(ipt, lay->L.ipt.data(), lay->L.ipt.size() * 8);
    if (oi) std::memcpy(oi, lay->L.OI.data(), lay->L.OI.size() * 8);
    return 0;
}

int mlr_layout_slots(const mlr_layout* lay, uint32_t cell, uint32_t* nb, uint32_t* info, double* nx, double* ny,
                     double* area, double* dxn) {
    if (!lay || cell >= lay->L.N) return 1;
    const mlr::Layout& L = lay->L;
    const uint32_t j = L.iperm[cell];
    for (uint32_t l = 0; l < L.S; ++l) {
        const size_t o = (size_t)l * L.N + j;
        if (nb) nb[l] = L.perm[L.nb[o]];
        if (info) info[l] = L.info[o];
        if (nx) nx[l] = L.nx[o];
        if (ny) ny[l] = L.ny[o];
        if (area) area[l] = L.area[o];
        if (dxn) dxn[l] = L.dxn[o];
    }
    return 0;
}

int mlr_probe_math(int which, const double* x, const double* y, double* out, int n) {
    return guarded(nullptr, [&] {
        if (n <= 0) return;
        double *dx = nullptr, *dy = nullptr, *dout = nullptr;
        auto chk = [](hipError_t e) { if (e != hipSuccess) mlr::fail(hipGetErrorString(e)); };
        chk(hipMalloc(&dx, n * 8));
        chk(hipMalloc(&dy, n * 8));
        chk(hipMalloc(&dout, n * 8));
        chk(hipMemcpy(dx, x, n * 8, hipMemcpyHostToDevice));
        chk(hipMemcpy(dy, y ? y : x, n * 8, hipMemcpyHostToDevice));
        chk(mlr::dev::probe_math(dx, dy, dout, n, which, nullptr));
        chk(hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost));
        (void)hipFree(dx);
        (void)hipFree(dy);
        (void)hipFree(dout);
    });
}

int mlr_probe_stream(int kind, int64_t bytes, int reps, double* ms_avg) {
    if (!ms_avg || bytes < 4096 || reps < 1 || kind < 0 || kind > 4) return 1;
    return guarded(nullptr, [&] {
        struct Buf {
            void* p = nullptr;
            ~Buf() { if (p) (void)hipFree(p); }
        } src, dst;
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev() { if (e) (void)hipEventDestroy(e); }
        } e0, e1;
        auto chk = [](hipError_t e) { if (e != hipSuccess) mlr::fail(hipGetErrorString(e)); };
        const size_t n = (size_t)bytes / 16 * 16;
        const size_t dbytes = (kind == 0 || kind == 2) ? n : (size_t)mlr::dev::stream_probe_threads(n) * 8;
        chk(hipMalloc(&src.p, n));
        chk(hipMalloc(&dst.p, dbytes));
        chk(hipMemset(src.p, 0, n));
        chk(hipEventCreate(&e0.e));
        chk(hipEventCreate(&e1.e));
        chk(mlr::dev::stream_probe(kind, src.p, dst.p, n, nullptr));  // warm-up
        chk(hipEventRecord(e0.e, nullptr));
        for (int r = 0; r < reps; ++r) chk(mlr::dev::stream_probe(kind, src.p, dst.p, n, nullptr));
        chk(hipEventRecord(e1.e, nullptr));
        chk(hipEventSynchronize(e1.e));
        float ms = 0.0f;
        chk(hipEventElapsedTime(&ms, e0.e, e1.e));
        *ms_avg = ms / reps;
    });
}

// ---------------------------------------------------------------- input parsing (host only)
int mlr_config_parse(const char* toml_path, mlr_mesh_gen* mesh, mlr_numerics_desc* num, double* run,
                     int32_t* init_type) {
    return guarded(nullptr, [&] {
        const mlr::Config c = mlr::config_from_toml(mlr::toml::parse_file(toml_path ? toml_path : ""));
        desc_from_config(c, mesh, num);
        if (run) {
            run[0] = c.use_cfl ? c.cfl : 0.0;
            run[1] = c.dt;
            run[2] = (double)c.n_steps;
            run[3] = c.t_stop;
            run[4] = c.t_wall_stop;
            run[5] = (double)c.check_interval;
            run[6] = c.check_nan ? 1.0 : 0.0;
        }
        if (init_type) *init_type = (int32_t)c.init.kind;
    });
}

int mlr_expr_eval(const char* expr, int64_t n, const double* x, const double* y, double* out) {
    if (!expr || n < 0 || (n > 0 && (!x || !y || !out))) return 1;
    return guarded(nullptr, [&] {
        const mlr::Expression e(expr);
        for (int64_t i = 0; i < n; ++i) out[i] = e.eval(x[i], y[i]);
    });
}

// ---------------------------------------------------------------- ensembles
int mlr_ensemble_create(const mlr_mesh_gen* gens, int n_cases, const mlr_numerics_desc* num, mlr_ensemble** out) {
    if (!out || !gens || n_cases <= 0) return 1;
    *out = nullptr;
    auto ens = std::make_unique<mlr_ensemble>();
    const int rc = guarded(nullptr, [&] {
        const mlr::Config c = config_from_desc(num);
        std::vector<mlr::MeshSpec> ms((size_t)n_cases);
        for (int k = 0; k < n_cases; ++k) {
            const mlr_mesh_gen& g = gens[k];
            if (g.type == MLR_MESH_CARTESIAN) ms[k].type = mlr::MeshType::Cartesian;
            else if (g.type == MLR_MESH_CARTESIAN_TRI) ms[k].type = mlr::MeshType::CartesianTri;
            else if (g.type == MLR_MESH_WEDGE) ms[k].type = mlr::MeshType::Wedge;
            else mlr::fail("MeshType::FILE not implemented.");
            ms[k].nx = g.nx;
            ms[k].ny = g.ny;
            ms[k].Lx = g.Lx;
            ms[k].Ly = g.Ly;
            ms[k].jitter = g.jitter;
            ms[k].seed = g.seed;
        }
        ens->e = std::make_unique<mlr::Ensemble>(c, ms);
    });
    if (rc == 0) *out = ens.release();
    return rc;
}

int mlr_ensemble_destroy(mlr_ensemble* ens) {
    delete ens;
    return 0;
}

int mlr_ensemble_sizes(const mlr_ensemble* ens, int64_t* n_cells) {
    if (!ens || !n_cells) return 1;
    for (int k = 0; k < ens->e->size(); ++k) n_cells[k] = ens->e->solver(k).mesh().n_cells;
    return 0;
}

int mlr_ensemble_init_quadrant(mlr_ensemble* ens, int k, double xs, double ys, const double* states) {
    if (!ens || !states || k < 0 || k >= ens->e->size()) return 1;
    return guarded(nullptr, [&] { quadrant_state(ens->e->solver(k), xs, ys, states); });
}

int mlr_ensemble_get_state(mlr_ensemble* ens, int k, double* U, double* prim) {
    if (!ens || k < 0 || k >= ens->e->size()) return 1;
    return guarded(nullptr, [&] { ens->e->solver(k).get_state(U, prim); });
}

int mlr_ensemble_get_volumes(mlr_ensemble* ens, int k, double* vol) {
    if (!ens || !vol || k < 0 || k >= ens->e->size()) return 1;
    return guarded(nullptr, [&] { ens->e->solver(k).get_volumes(vol); });
}

int mlr_ensemble_get_time(mlr_ensemble* ens, int k, double* t, int64_t* step) {
    if (!ens || k < 0 || k >= ens->e->size()) return 1;
    if (t) *t = ens->e->solver(k).time();
    if (step) *step = ens->e->solver(k).step();
    return 0;
}

int mlr_ensemble_run_steps(mlr_ensemble* ens, int64_t n) {
    if (!ens) return 1;
    return guarded(nullptr, [&] { ens->e->run_steps(n); });
}

int mlr_ensemble_run_steps_profiled(mlr_ensemble* ens, int64_t n, double* stats) {
    if (!ens || !stats) return 1;
    return guarded(nullptr, [&] { ens->e->run_steps_profiled(n, stats); });
}
int mlr_ensemble_autotune(mlr_ensemble* ens, int32_t* tune) {
    if (!ens || !tune) return 1;
    return guarded(nullptr, [&] {
        ens->e->autotune();
        tune_out(ens->e->tune(), tune);
    });
}
int mlr_ensemble_set_tune(mlr_ensemble* ens, const int32_t* tune) {
    if (!ens || !tune) return 1;
    return guarded(nullptr, [&] { ens->e->set_tune(tune_in(tune)); });
}

int mlr_ensemble_set_indicator(mlr_ensemble* ens, const float* params, double threshold) {
    if (!ens) return 1;
    return guarded(nullptr, [&] { ens->e->set_indicator(params, threshold); });
}

int mlr_ensemble_gated(mlr_ensemble* ens, int reset, int64_t* gated) {
    if (!ens || !gated) return 1;
    return guarded(nullptr, [&] {
        int64_t g = 0;
        for (int k = 0; k < ens->e->size(); ++k) g += (int64_t)ens->e->solver(k).gated_cells(reset != 0);
        *gated = g;
    });
}

// ---------------------------------------------------------------- output
int mlr_write_vtu(mlr_ctx* ctx, const char* path, const char* const* variables, int n_vars) {
    if (!ctx || !path || !variables || n_vars <= 0) return 1;
    return guarded(ctx, [&] {
        const mlr::Mesh& m = ctx->solver.mesh();
        const size_t N = m.n_cells;
        std::vector<double> U(4 * N), W(5 * N), cfl(N);
        ctx->solver.get_state(U.data(), W.data());
        ctx->solver.get_cfl_local(cfl.data());
        std::vector<std::string> v(variables, variables + n_vars);
        mlr::write_vtu(path, m, v, U.data(), W.data(), cfl.data());
    });
}

int mlr_layout_write_vtu(const mlr_layout* lay, const char* path, const char* const* variables, int n_vars,
                         const double* U, const double* prim, const double* cfl) {
    if (!lay || !path || !variables || n_vars <= 0 || !U || !prim || !cfl) return 1;
    return guarded(nullptr, [&] {
        std::vector<std::string> v(variables, variables + n_vars);
        mlr::write_vtu(path, lay->mesh, v, U, prim, cfl);
    });
}

// ---------------------------------------------------------------- learned indicator
int mlr_dataset_create(int64_t capacity, mlr_dataset** out) {
    if (!out || capacity <= 0) return 1;
    *out = nullptr;
    return guarded(nullptr, [&] { *out = new mlr_dataset((uint64_t)capacity); });
}

int mlr_dataset_destroy(mlr_dataset* ds) {
    delete ds;
    return 0;
}

int mlr_dataset_size(const mlr_dataset* ds, int64_t* n) {
    if (!ds || !n) return 1;
    *n = (int64_t)ds->ds.n;
    return 0;
}

int mlr_dataset_get(const mlr_dataset* ds, int64_t start, int64_t n, uint16_t* X, float* y) {
    if (!ds || start < 0 || n < 0) return 1;
    return guarded(nullptr, [&] { ds->ds.get((uint64_t)start, (uint64_t)n, X, y); });
}

int mlr_collect_features(mlr_ctx* ctx, mlr_dataset* ds) {
    if (!ctx || !ds) return 1;
    return guarded(ctx, [&] { ctx->solver.collect_features(ds->ds); });
}

int mlr_indicator_features(mlr_ctx* ctx, uint16_t* X, float* y) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.indicator_features(X, y); });
}

int mlr_mlp_forward(const float* params, const uint16_t* X, int64_t n, float* z) {
    if (!params || (n > 0 && (!X || !z)) || n < 0) return 1;
    return guarded(nullptr, [&] { mlr::mlp_forward(params, X, (uint64_t)n, z); });
}

int mlr_mlp_grad(const float* params, const mlr_dataset* ds, const uint32_t* idx, int64_t B, float pos_w,
                 float neg_w, float* grad, float* loss) {
    if (!params || !ds || !idx || !grad || !loss || B <= 0 || B > 0xFFFFFFFFll) return 1;
    return guarded(nullptr, [&] { mlr::mlp_grad(params, ds->ds, idx, (uint32_t)B, pos_w, neg_w, grad, loss); });
}

int mlr_mlp_train(float* params, const mlr_dataset* ds, const uint32_t* pool, int64_t n_pool, int steps, int batch,
                  float lr, uint64_t seed, float pos_w, float neg_w, float* loss_hist) {
    if (!params || !ds || !pool || n_pool <= 0 || steps <= 0 || batch <= 0) return 1;
    return guarded(nullptr, [&] {
        mlr::mlp_train(params, ds->ds, pool, (uint64_t)n_pool, steps, (uint32_t)batch, lr, seed, pos_w, neg_w,
                       loss_hist);
    });
}

int mlr_set_indicator(mlr_ctx* ctx, const float* params, double threshold) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.set_indicator(params, threshold); });
}

int mlr_indicator_cadence(mlr_ctx* ctx, int every_stage) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.set_indicator_cadence(every_stage != 0); });
}

int mlr_indicator_split(mlr_ctx* ctx, int split) {
    if (!ctx) return 1;
    return guarded(ctx, [&] { ctx->solver.set_indicator_split(split != 0); });
}

int mlr_ensemble_indicator_split(mlr_ensemble* ens, int split) {
    if (!ens) return 1;
    return guarded(nullptr, [&] { ens->e->set_indicator_split(split != 0); });
}

int mlr_ensemble_set_persistent(mlr_ensemble* ens, int on) {
    if (!ens) return 1;
    return guarded(nullptr, [&] { ens->e->set_persistent(on != 0); });
}
int mlr_ensemble_persistent_wait_limit(mlr_ensemble* ens, double ms) {
    if (!ens) return 1;
    return guarded(nullptr, [&] { ens->e->set_persistent_wait_limit(ms); });
}
int mlr_ensemble_persistent_grid(const mlr_ensemble* ens, int32_t* blocks) {
    if (!ens || !blocks) return 1;
    *blocks = ens->e->persistent_grid();
    return 0;
}

int mlr_indicator_decisions(mlr_ctx* ctx, uint8_t* out) {
    if (!ctx || !out) return 1;
    return guarded(ctx, [&] { ctx->solver.get_gate_decisions(out); });
}

int mlr_ensemble_indicator_cadence(mlr_ensemble* ens, int every_stage) {
    if (!ens) return 1;
    return guarded(nullptr, [&] { ens->e->set_indicator_cadence(every_stage != 0); });
}

int mlr_indicator_gated(mlr_ctx* ctx, int reset, int64_t* gated) {
    if (!ctx || !gated) return 1;
    return guarded(ctx, [&] { *gated = (int64_t)ctx->solver.gated_cells(reset != 0); });
}

int mlr_get_volumes(const mlr_ctx* ctx, double* vol) {
    if (!ctx || !vol) return 1;
    ctx->solver.get_volumes(vol);
    return 0;
}

int mlr_host_pow(const double* x, const double* y, double* out, int n) {
    if (n < 0 || (n > 0 && (!x || !y || !out))) return 1;
    for (int i = 0; i < n; ++i) out[i] = mlr::ddpow::pow_dd(x[i], y[i]);
    return 0;
}

int mlr_host_pow6(const double* x, double* out, int n) {
    if (n < 0 || (n > 0 && (!x || !out))) return 1;
    for (int i = 0; i < n; ++i) out[i] = mlr::pow6_cr(x[i]);
    return 0;
}

}  // extern "C"
