#include "solver.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>

#include "expr.hpp"
#include "io.hpp"
#include "learn.hpp"

namespace mlr {

namespace {

#define HIP_OK(x)                                                                          \
    do {                                                                                   \
        const hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) fail(std::string("HIP error: ") + hipGetErrorString(e_) +    \
                                   " at " + __FILE__ + ":" + std::to_string(__LINE__));    \
    } while (0)

double wall_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <class T>
T lookup_enum(const std::string& key, std::initializer_list<std::pair<const char*, T>> table,
              const std::string& what) {
    for (auto& kv : table)
        if (key == kv.first) return kv.second;
    fail("Unknown " + what + ": " + key + ".");
}

// physics.h:852-867 on the host (same operation order as the device)
void host_prims(const dev::Physics& ph, const double* U, double* W) {
    const double rho = U[0];
    const double u0 = U[1] / rho, u1 = U[2] / rho;
    const double E = U[3] / rho;
    const double e = E - 0.5 * (u0 * u0 + u1 * u1);
    const double p = std::fmax(ph.p_min, std::fmin(ph.p_max, (ph.gamma - 1.0) * rho * e));
    W[0] = u0; W[1] = u1; W[2] = p; W[3] = e / ph.cv; W[4] = e + p / rho;
}

constexpr int64_t kDtHistCap = 1 << 20;

}  // namespace

// ------------------------------------------------------------ input
Config config_from_toml(const toml::Value& in) {
    using namespace toml;
    Config c;
    // [mesh] (mesh.cpp:32-66; jitter/seed are this engine's extension keys)
    const std::string mt = find_or_str(in, {"mesh", "type"}, "file");
    c.mesh_type = lookup_enum<MeshType>(mt, {{"file", MeshType::File}, {"cartesian", MeshType::Cartesian},
                                             {"cartesian_tri", MeshType::CartesianTri}, {"wedge", MeshType::Wedge}},
                                        "mesh type");
    if (c.mesh_type == MeshType::File) fail("MeshType::FILE not implemented.");
    c.nx = (uint32_t)find_or_int(in, {"mesh", "Nx"}, 100);
    c.ny = (uint32_t)find_or_int(in, {"mesh", "Ny"}, 100);
    c.Lx = find_or(in, {"mesh", "Lx"}, 1.0);
    c.Ly = find_or(in, {"mesh", "Ly"}, 1.0);
    c.jitter = find_or(in, {"mesh", "jitter"}, 0.0);
    c.seed = (uint64_t)find_or_int(in, {"mesh", "seed"}, 0);

    // [physics] (solver.cpp:88-108, physics.cpp:27-53)
    const std::string pt = find_or_str(in, {"physics", "type"}, "euler");
    if (pt != "euler" && pt != "navier-stokes") fail("Unknown physics type: " + pt + ".");
    const Value& ph = find(in, "physics");
    for (const char* k : {"gamma", "p_ref", "T_ref", "rho_ref"})
        if (!ph.contains(k)) fail(std::string("Missing ") + k + " for physics: " + pt + ".");
    c.gamma = find(ph, "gamma").as_float();
    c.p_ref = find(ph, "p_ref").as_float();
    c.T_ref = find(ph, "T_ref").as_float();
    c.rho_ref = find(ph, "rho_ref").as_float();
    c.p_min = find_or(in, {"physics", "p_min"}, -1e20);
    c.p_max = find_or(in, {"physics", "p_max"}, 1e20);
    if (pt == "navier-stokes") {
        if (!ph.contains("mu_ref")) fail("Missing mu_ref for physics: navier-stokes.");
        c.viscous = true;
        c.mu_ref = find(ph, "mu_ref").as_float();
        c.T_mu = find_or(in, {"physics", "T_mu_ref"}, c.T_ref);
        c.S_mu = find_or(in, {"physics", "sutherland_S"}, -1.0);
        c.Pr = find_or(in, {"physics", "Pr"}, 0.72);
        if (!(c.mu_ref >= 0.0) || !(c.T_mu > 0.0) || !(c.Pr > 0.0)) fail("Invalid viscosity parameters.");
    }

    // [numerics] (solver.cpp:110-186, face_reconstruction.cpp:109-152)
    const Value& fr = find(in, "numerics", "face_reconstruction");
    c.recon = lookup_enum<Recon>(find_or_str(fr, {"type"}, "FO"), {{"FO", Recon::FO}, {"TENO", Recon::TENO}},
                                 "face reconstruction type");
    c.riemann = lookup_enum<Riemann>(find_or_str(in, {"numerics", "riemann_solver"}, "HLLC"),
                                     {{"Rusanov", Riemann::Rusanov}, {"HLL", Riemann::HLL},
                                      {"HLLC", Riemann::HLLC}, {"Roe", Riemann::Roe}},
                                     "Riemann solver type");
    c.integrator = lookup_enum<Integrator>(find_or_str(in, {"numerics", "time_integrator"}, "LSSSPRK3"),
                                           {{"FE", Integrator::FE}, {"RK4", Integrator::RK4},
                                            {"SSPRK3", Integrator::SSPRK3}},
                                           "time integrator type");
    c.check_nan = find_or_bool(in, {"numerics", "check_nan"}, false);
    // extension key (no reference counterpart): the positivity-preserving flux limiter
    c.positivity = find_or_bool(in, {"numerics", "positivity_limiter"}, false);
    if (c.recon == Recon::TENO) {
        c.teno.basis = lookup_enum<BasisType>(find_or_str(fr, {"basis_type"}, "monomial"),
                                              {{"monomial", BasisType::Monomial}, {"legendre", BasisType::Legendre}},
                                              "basis type");
        c.teno.order = (int)(uint8_t)find(fr, "basis_order").as_int();
        c.teno.factor = find_or(fr, {"max_stencil_size_factor"}, 2.0);
        const std::string qtc = find_or_str(fr, {"quadrature_type_cell"}, "triangle_dunavant");
        const std::string qtf = find_or_str(fr, {"quadrature_type_face"}, "gauss_legendre");
        if (qtc != "triangle_dunavant") fail("Unknown quadrature type: " + qtc);
        if (qtf != "gauss_legendre") fail("Unknown quadrature type: " + qtf);
        c.teno.quad_order_cell = (int)(uint8_t)find_or_int(fr, {"quadrature_order_cell"}, c.teno.order + 1);
        c.teno.quad_order_face = (int)(uint8_t)find_or_int(fr, {"quadrature_order_face"}, (c.teno.order + 1) / 2);
        c.teno.variant = lookup_enum<TenoVariant>(find_or_str(fr, {"teno_variant"}, "reference"),
                                                  {{"reference", TenoVariant::Reference},
                                                   {"corrected", TenoVariant::Corrected}},
                                                  "TENO variant");
    }

    // [[boundaries]] (solver.cpp:188-237, boundary_upt.cpp:40-71, boundary_p_out.cpp:36-52)
    const Value& bl = find(in, "boundaries");
    if (bl.kind != Value::Array) fail("boundaries must be an array of tables");
    for (const Value& b : bl.arr) {
        if (!b.contains("name")) fail("Boundary name not specified.");
        if (!b.contains("type")) fail("Boundary type not specified.");
        BoundarySpec s;
        s.zone = find(b, "name").as_string();
        const std::string ty = find(b, "type").as_string();
        s.type = lookup_enum<BCType>(ty, {{"symmetry", BCType::Symmetry}, {"extrapolation", BCType::Extrapolation},
                                          {"wall_adiabatic", BCType::WallAdiabatic}, {"upt", BCType::UPT},
                                          {"p_out", BCType::POut}, {"wall_noslip", BCType::WallNoSlip}},
                                     "boundary type");
        if (s.type == BCType::UPT) {
            for (const char* k : {"u", "p", "T"})
                if (!b.contains(k)) fail(std::string("Missing ") + k + " for boundary: " + s.zone + ".");
            const Value& u = find(b, "u");
            if (u.kind != Value::Array || u.arr.size() != 2) fail("Invalid u for boundary: " + s.zone + ".");
            s.u[0] = u.arr[0].as_float();
            s.u[1] = u.arr[1].as_float();
            s.p = find(b, "p").as_float();
            s.T = find(b, "T").as_float();
        } else if (s.type == BCType::POut) {
            if (!b.contains("p")) fail("Missing p for boundary: " + s.zone + ".");
            s.p = find(b, "p").as_float();
        }
        c.bcs.push_back(s);
    }

    // [run] (solver.cpp:239-273)
    const Value* run = in.get("run");
    if (!run) fail("Run parameters not specified.");
    const bool has_dt = run->contains("dt"), has_cfl = run->contains("cfl");
    if (!has_dt && !has_cfl) fail("Either dt or cfl must be specified.");
    if (has_dt && has_cfl) fail("Only one of dt or cfl can be specified.");
    if (!run->contains("n_steps") && !run->contains("t_stop") && !run->contains("t_wall_stop"))
        fail("Either n_steps, t_stop, or t_wall_stop must be specified.");
    c.use_cfl = has_cfl;
    if (has_dt) c.dt = find(*run, "dt").as_float();
    else c.cfl = find(*run, "cfl").as_float();
    c.n_steps = run->contains("n_steps") ? (int64_t)(uint32_t)find(*run, "n_steps").as_int() : (int64_t)UINT32_MAX;
    c.t_stop = find_or(in, {"run", "t_stop"}, -1.0);
    c.t_wall_stop = find_or(in, {"run", "t_wall_stop"}, -1.0);
    c.check_interval = (uint32_t)find_or_int(in, {"output", "check_interval"}, 1);

    // [[write_data]] (solver.cpp:283-295, data_writer.cpp:29-77)
    if (const Value* wd = in.get("write_data")) {
        if (wd->kind != Value::Array) fail("write_data must be an array of tables.");
        for (const Value& o : wd->arr) {
            for (const char* k : {"prefix", "format", "interval", "variables"})
                if (!o.contains(k)) fail(std::string("DataWriter: ") + k + " not specified.");
            WriteSpec w;
            w.prefix = find(o, "prefix").as_string();
            w.interval = (uint32_t)find(o, "interval").as_int();
            const std::string fmt = find(o, "format").as_string();
            if (fmt == "tecplot") fail("DataWriter::write_tecplot not implemented.");
            if (fmt != "vtu") fail("DataWriter: Unknown format type: " + fmt + ".");
            const Value& vars = find(o, "variables");
            if (vars.kind != Value::Array || vars.arr.empty()) fail("DataWriter: No variables specified.");
            for (const Value& v : vars.arr) {
                if (!is_output_variable(v.as_string())) fail("DataWriter: Unknown variable: " + v.as_string() + ".");
                w.variables.push_back(v.as_string());
            }
            if (w.interval == 0) fail("DataWriter: interval must be positive.");
            c.writers.push_back(w);
        }
    }

    // [initialize] (solver_initialize.cpp:35-186; "quadrant" is this engine's
    // shorthand for the four-state Riemann problems)
    const std::string it = find_or_str(in, {"initialize", "type"}, "constant");
    const Value* ini = in.get("initialize");
    if (it == "constant") {
        c.init.kind = InitSpec::Constant;
        if (!ini || !ini->contains("u")) fail("Missing u for initialization: constant.");
        const Value& u = find(*ini, "u");
        if (u.kind != Value::Array || u.arr.size() != 2) fail("u must be a 2-element array for initialization: constant.");
        if (!ini->contains("p")) fail("Missing p for initialization: constant.");
        if (!ini->contains("T")) fail("Missing T for initialization: constant.");
        c.init.u[0] = u.arr[0].as_float();
        c.init.u[1] = u.arr[1].as_float();
        c.init.p = find(*ini, "p").as_float();
        c.init.T = find(*ini, "T").as_float();
    } else if (it == "analytical") {
        c.init.kind = InitSpec::Analytical;
        if (!ini || !ini->contains("u")) fail("Missing u for initialization: analytical.");
        const Value& u = find(*ini, "u");
        if (u.kind != Value::Array || u.arr.size() != 2) fail("u must be a 2-element array for initialization: analytical.");
        const int n = (int)ini->contains("rho") + (int)ini->contains("p") + (int)ini->contains("T");
        if (n != 2) fail("Exactly two of rho, p, and T must be specified for initialization: analytical.");
        c.init.u_expr[0] = u.arr[0].as_string();
        c.init.u_expr[1] = u.arr[1].as_string();
        c.init.rho_expr = find_or_str(*ini, {"rho"}, "");
        c.init.p_expr = find_or_str(*ini, {"p"}, "");
        c.init.T_expr = find_or_str(*ini, {"T"}, "");
    } else if (it == "quadrant") {
        c.init.kind = InitSpec::Quadrant;
        const Value& sp = find(*ini, "split");
        c.init.xs = sp.arr.at(0).as_float();
        c.init.ys = sp.arr.at(1).as_float();
        const char* names[4] = {"UR", "UL", "LL", "LR"};
        for (int q = 0; q < 4; ++q) {
            const Value& st = find(*ini, names[q]);
            if (st.kind != Value::Array || st.arr.size() != 4) fail(std::string("quadrant state ") + names[q] + " needs [rho, u, v, p]");
            for (int k = 0; k < 4; ++k) c.init.states[4 * q + k] = st.arr[k].as_float();
        }
    } else {
        fail("Unknown initialization type: " + it + ".");
    }
    return c;
}

// ------------------------------------------------------------ solver
Solver::Solver() = default;

Solver::~Solver() { free_device(); }

void Solver::free_device() {
    if (step_exec_) { (void)hipGraphExecDestroy(step_exec_); step_exec_ = nullptr; }
    if (step_graph_) { (void)hipGraphDestroy(step_graph_); step_graph_ = nullptr; }
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    dev_bytes_ = 0;
    net_buf_ = nullptr;
    theta_ = nullptr;
    gate_ = false;
    for (auto& e : prof_) (void)hipEventDestroy(e.second);
    prof_.clear();
    prof_used_ = 0;
    if (stream_) { (void)hipStreamDestroy(stream_); stream_ = nullptr; }
}

int Solver::init(const std::string& input_file) {
    const toml::Value in = toml::parse_file(input_file);
    init(config_from_toml(in));
    return 0;
}

void Solver::init(const Config& cfg) {
    cfg_ = cfg;
    Mesh m;
    switch (cfg.mesh_type) {
    case MeshType::Cartesian: m = make_cartesian(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.jitter, cfg.seed); break;
    case MeshType::CartesianTri: m = make_cartesian_tri(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly, cfg.jitter, cfg.seed); break;
    case MeshType::Wedge: m = make_wedge(cfg.nx, cfg.ny, cfg.Lx, cfg.Ly); break;
    default: fail("MeshType::FILE not implemented.");
    }
    setup(std::move(m));
}

void Solver::init(const Config& cfg, Mesh&& mesh) {
    cfg_ = cfg;
    setup(std::move(mesh));
}

void Solver::setup(Mesh&& mesh) {
    free_device();
    mesh_ = std::move(mesh);
    // physics.cpp:69-73
    ph_.gamma = cfg_.gamma;
    ph_.R = cfg_.p_ref / (cfg_.T_ref * cfg_.rho_ref);
    ph_.cp = ph_.R * cfg_.gamma / (cfg_.gamma - 1.0);
    ph_.cv = ph_.cp / cfg_.gamma;
    ph_.p_min = cfg_.p_min;
    ph_.p_max = cfg_.p_max;
    ph_.viscous = cfg_.viscous ? 1 : 0;
    ph_.mu_ref = cfg_.mu_ref;
    ph_.T_mu = cfg_.T_mu;
    ph_.S_mu = cfg_.S_mu;
    ph_.Pr = cfg_.Pr;
    if (cfg_.integrator != Integrator::FE && cfg_.integrator != Integrator::RK4 &&
        cfg_.integrator != Integrator::SSPRK3)
        fail("Unknown time integrator type.");
    if (cfg_.positivity && cfg_.integrator == Integrator::RK4)
        fail("the positivity limiter needs forward Euler or SSP-RK3 (RK4 is no convex combination of Euler steps)");
    if (cfg_.positivity && cfg_.viscous) fail("the positivity limiter is for the Euler equations");
    L_ = build_layout(mesh_, cfg_.recon, cfg_.teno, cfg_.bcs, cfg_.n_threads, /*allow_device=*/true);
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    upload();
    t_ = 0.0;
    step_ = 0;
    wall0_ = wall_seconds();
    init_solution();
}

void Solver::precompute_aplus() {
    struct Tmp {
        std::vector<void*> p;
        ~Tmp() { for (void* q : p) (void)hipFree(q); }
        void* get(size_t bytes) {
            void* q = nullptr;
            HIP_OK(hipMalloc(&q, bytes ? bytes : 8));
            p.push_back(q);
            return q;
        }
    } tmp;
    const uint32_t N = L_.N;
    std::vector<uint32_t> tri(3 * (size_t)N);
    for (uint32_t j = 0; j < N; ++j)
        for (int l = 0; l < 3; ++l) tri[3 * (size_t)j + l] = mesh_.node_of_cell(L_.perm[j], l);
    dev::PreInput in{};
    in.N = N;
    void* xy = tmp.get(mesh_.node_coords.size() * sizeof(double));
    void* tr = tmp.get(tri.size() * sizeof(uint32_t));
    void* ids = tmp.get(L_.st_ids.size() * sizeof(int32_t));
    int* err = (int*)tmp.get(sizeof(int));
    HIP_OK(hipMemcpy(xy, mesh_.node_coords.data(), mesh_.node_coords.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(tr, tri.data(), tri.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ids, L_.st_ids.data(), L_.st_ids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(err, 0, sizeof(int)));
    in.xy = (const double*)xy;
    in.tri = (const uint32_t*)tr;
    in.ids = (const int32_t*)ids;
    in.info = info_;
    in.nq = (int32_t)L_.qcell_w.size();
    if (in.nq > 8 || L_.K > 10) fail("device precompute: quadrature or basis too large");
    for (int q = 0; q < in.nq; ++q) {
        in.qp[2 * q] = L_.qcell_pts[2 * (size_t)q];
        in.qp[2 * q + 1] = L_.qcell_pts[2 * (size_t)q + 1];
        in.qw[q] = L_.qcell_w[q];
    }
    for (int k = 0; k < 2 * L_.K; ++k) in.poly[k] = L_.poly[k];
    for (int k = 0; k < L_.K; ++k) in.ipt[k] = L_.ipt[k];
    HIP_OK(dev::teno_aplus(in, L_.K, L_.M, A_, ta_, err, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    int e = 0;
    HIP_OK(hipMemcpy(&e, err, sizeof(int), hipMemcpyDeviceToHost));
    if (e & 1) fail("Matrix is singular.");
    if (e & 2) fail("a TENO stencil needs the full-matrix pseudo-inverse branch, which this layout does not store");
}

void Solver::upload() {
    auto dalloc = [&](size_t bytes) {
        void* p = nullptr;
        HIP_OK(hipMalloc(&p, bytes ? bytes : 8));
        allocs_.push_back(p);
        dev_bytes_ += bytes;
        return p;
    };
    auto up = [&](auto*& dst, const auto& v) {
        using T = std::remove_reference_t<decltype(v[0])>;
        dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(dalloc(v.size() * sizeof(T)));
        if (!v.empty()) HIP_OK(hipMemcpy(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    };
    const size_t N = L_.N;
    up(volume_, L_.volume);
    up(nb_, L_.nb);
    up(info_, L_.info);
    up(nx_, L_.nx);
    up(ny_, L_.ny);
    up(area_, L_.area);
    up(dxn_, L_.dxn);
    up(qw_, L_.qface_w);
    if (cfg_.viscous) {
        up(dcx_, L_.dcx);
        up(dcy_, L_.dcy);
        up(mfx_, L_.mfx);
        up(mfy_, L_.mfy);
    }
    std::vector<dev::BCParam> bct(std::max<size_t>(cfg_.bcs.size(), 1));
    for (size_t b = 0; b < cfg_.bcs.size(); ++b) {
        const BoundarySpec& s = cfg_.bcs[b];
        dev::BCParam& p = bct[b];
        p.type = (int32_t)s.type;
        p.u0 = s.u[0];
        p.u1 = s.u[1];
        p.p = s.p;
        p.T = s.T;
        if (s.type == BCType::UPT) {  // boundary_upt.cpp:58-60 (host EOS)
            p.rho = s.p / (ph_.R * s.T);
            p.e = ph_.cv * s.T;
            p.h = p.e + s.p / p.rho;
        }
    }
    up(bc_, bct);
    if (L_.teno) {
        if (!L_.st_loc8.empty()) up(loc8_, L_.st_loc8);
        else up(loc16_, L_.st_loc16);
        up(blk_off_, L_.blk_off);
        up(blk_ids_, L_.blk_ids);
        if (L_.device_aplus) {
            // pseudo-inverses and member areas built on the GPU from the
            // stencil ids, the node coordinates and the cells' nodes
            const size_t Np = L_.n_pad(), Mp = (size_t)L_.M - 1, Kp = (size_t)L_.K - 1;
            A_ = (double*)dalloc((size_t)L_.G * Mp * Kp * Np * sizeof(double));
            ta_ = (double*)dalloc((size_t)L_.G * Mp * Np * sizeof(double));
            precompute_aplus();
        } else {
            up(ta_, L_.st_ta);
            up(A_, L_.st_A);
        }
        up(fbasis_, L_.fbasis);
        if (L_.S == 3 && L_.Q == 1 && L_.K == 6 && L_.M == 12 && !L_.st_loc8.empty() &&
            L_.tspec.basis == BasisType::Legendre && L_.qface_pts.size() == 1) {
            // the cells' node coordinates, for the plain TENO pass's derived face basis
            std::vector<double> cv(6 * (size_t)L_.N);
            for (uint32_t j = 0; j < L_.N; ++j)
                for (int r = 0; r < 3; ++r) {
                    const uint32_t nd = mesh_.node_of_cell(L_.perm[j], r);
                    cv[6 * (size_t)j + 2 * r] = mesh_.nc(nd, 0);
                    cv[6 * (size_t)j + 2 * r + 1] = mesh_.nc(nd, 1);
                }
            up(cvert_, cv);
        }
        // host copies of the big tables are not needed any more
        std::vector<int32_t>().swap(L_.st_ids);
        std::vector<uint8_t>().swap(L_.st_loc8);
        std::vector<uint16_t>().swap(L_.st_loc16);
        std::vector<int32_t>().swap(L_.blk_ids);
        std::vector<double>().swap(L_.st_ta);
        std::vector<double>().swap(L_.st_A);
        std::vector<double>().swap(L_.fbasis);
    }
    auto dz = [&](size_t n) {
        double* p = (double*)dalloc(n * sizeof(double));
        HIP_OK(hipMemset(p, 0, n * sizeof(double)));
        return p;
    };
    Un_ = dz(4 * N);
    Ua_ = dz(4 * N);
    if (cfg_.viscous) {
        vp_ = dz(4 * N);
        vg_ = dz(6 * N);
    }
    Ub_ = dz(4 * N);
    W_ = dz(4 * N);
    prims_ = dz(4 * N);
    h_ = dz(N);
    rhs_ = dz(4 * N);
    lam_cell_ = dz(N);
    fs_ = L_.teno ? dz(4 * (size_t)L_.S * L_.Q * N) : dz(1);
    dt_dev_ = dz(1);
    dt_hist_ = dz(kDtHistCap);
    dt_hist_cap_ = kDtHistCap;
    lam_blk_ = (unsigned long long*)dz((N + 63) / 64);
    step_dev_ = (int64_t*)dz(1);
    gated_slots_ = 4 * (size_t)((N + 63) / 64);
    gated_ = (uint32_t*)dz((gated_slots_ + 1) / 2);
    dec_ = (uint8_t*)dalloc(N);
    HIP_OK(hipMemset(dec_, 0, N));
}

dev::Geom Solver::geom() const {
    dev::Geom g;
    g.N = L_.N;
    g.S = L_.S;
    g.Q = L_.Q;
    g.volume = volume_;
    g.nb = nb_;
    g.info = info_;
    g.nx = nx_;
    g.ny = ny_;
    g.area = area_;
    g.dxn = dxn_;
    g.qw = qw_;
    g.bc = bc_;
    g.dcx = dcx_;
    g.dcy = dcy_;
    g.mfx = mfx_;
    g.mfy = mfy_;
    return g;
}

dev::Teno Solver::teno() const {
    dev::Teno t{};
    t.K = L_.K;
    t.M = L_.M;
    t.variant = (int32_t)L_.tspec.variant;
    t.basis0 = L_.basis0;
    t.loc8 = loc8_;
    t.loc16 = loc16_;
    t.blk_off = blk_off_;
    t.blk_ids = blk_ids_;
    t.ta = ta_;
    t.A = A_;
    t.fbasis = fbasis_;
    const int Kp = L_.K - 1;
    for (int i = 0; i < Kp; ++i)
        for (int j = 0; j < Kp; ++j) t.OI[i * Kp + j] = L_.OI[(size_t)(i + 1) * L_.K + (j + 1)];
    t.gamma = ph_.gamma;
    t.gated = gated_;
    t.gate_dec = gate_dec_;
    t.lin = lin_;
    t.blk_mixed = (gate_ && gate_split_) ? blk_mixed_ : nullptr;
    t.compact = compact_ ? 1 : 0;
    t.fbc = fbc_;
    t.gate = gate_ ? 1 : 0;
    // the derived face basis serves the plain pass (no gate, exact tables kept)
    t.derived = (tune_.teno_face_basis == 1 && cvert_ && !compact_ && !gate_) ? 1 : 0;
    t.cvert = cvert_;
    for (int k = 0; k < L_.K && k < 10; ++k) {
        t.ipt[k] = L_.ipt[k];
        t.poly[2 * k] = L_.poly[2 * k];
        t.poly[2 * k + 1] = L_.poly[2 * k + 1];
    }
    t.qf = L_.qface_pts.empty() ? 0.0 : L_.qface_pts[0];
    if (gate_) t.net = dev::net_view(net_buf_, gate_thr_, nullptr);
    return t;
}

void Solver::drop_graph() {
    if (step_exec_) { HIP_OK(hipGraphExecDestroy(step_exec_)); step_exec_ = nullptr; }
    if (step_graph_) { HIP_OK(hipGraphDestroy(step_graph_)); step_graph_ = nullptr; }
}

// The member areas folded into the pseudo-inverses (a = (A+ diag(ta)) (u_i -
// u_c) instead of A+ (ta (u_i - u_c))) and the face-point basis values taken
// from the reference triangle's edges instead of mapping each cell's face
// points through its inverse Jacobian: the same quantities in exact
// arithmetic, 472 fewer bytes per cell (the 11 x 4 member areas and the 3 x 5
// face-basis values) for the TENO pass to read, different rounding.
void Solver::compact_tables() {
    if (compact_) return;
    if (!L_.teno || L_.tspec.variant != TenoVariant::Corrected)
        fail("compact TENO tables: TENO with the corrected weighting only");
    if (!(L_.K == 6 && L_.M == 12 && L_.S == 3 && L_.Q == 1 && loc8_))
        fail("compact TENO tables: p = 2 with the default stencil size only");
    const int Kp = L_.K - 1, Q = (int)L_.Q;
    const uint64_t rows = (uint64_t)((L_.N + 63) / 64) * L_.G * (L_.M - 1);
    HIP_OK(dev::teno_fold_areas(A_, ta_, rows, Kp, stream_));
    // basis values at the face point (the midpoint) of each edge of the
    // reference triangle: e = 0 nodes 0-1, 1 nodes 1-2, 2 nodes 2-0
    std::vector<double> tab((size_t)3 * Kp, 0.0);
    const double R[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
    for (int e = 0; e < 3; ++e) {
        const int ia = e, ib = (e + 1) % 3;
        double xq[2];
        for (int i = 0; i < 2; i++) xq[i] = (L_.qface_pts[0] + 1.0) * 0.5 * (R[ib][i] - R[ia][i]) + R[ia][i];
        for (int k = 1; k < L_.K; ++k)
            tab[(size_t)e * Kp + (k - 1)] =
                basis_2d(L_.tspec.basis, L_.poly[2 * k], L_.poly[2 * k + 1], xq[0], xq[1]) - L_.ipt[k];
    }
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, tab.size() * sizeof(double)));
    allocs_.push_back(p);
    dev_bytes_ += tab.size() * sizeof(double);
    fbc_ = static_cast<double*>(p);
    HIP_OK(hipMemcpy(fbc_, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    synchronize();
    // the member areas and per-cell face basis are not read any more
    auto release = [&](double*& buf, uint64_t bytes) {
        for (void*& a : allocs_)
            if (a == buf) {
                HIP_OK(hipFree(a));
                a = nullptr;
            }
        dev_bytes_ -= bytes;
        buf = nullptr;
    };
    const size_t Np = L_.n_pad();
    release(ta_, (uint64_t)L_.G * (L_.M - 1) * Np * sizeof(double));
    release(fbasis_, (uint64_t)L_.S * Q * Kp * Np * sizeof(double));
    compact_ = true;
    drop_graph();
    tuned_ = false;
    if (lin_) {
        HIP_OK(dev::teno_linear_operator(teno(), info_, L_.N, (int)(L_.S * L_.Q), lin_, stream_));
        synchronize();
    }
}

void Solver::set_indicator(const float* params, double threshold) {
    if (!params) {
        gate_ = false;
        drop_graph();
        tuned_ = false;
        return;
    }
    if (!L_.teno) fail("the learned indicator gates TENO reconstruction only");
    if (L_.M < kIndFeatures / 4) fail("the learned indicator needs central stencils of at least 12 cells");
    if (!(threshold > 0.0 && threshold <= 1.0)) fail("indicator threshold must lie in (0, 1]");
    if (!net_buf_) {
        HIP_OK(hipMalloc(&net_buf_, dev::net_bytes()));
        allocs_.push_back(net_buf_);
        void* th = nullptr;
        HIP_OK(hipMalloc(&th, kIndParams * sizeof(float)));
        allocs_.push_back(th);
        theta_ = static_cast<float*>(th);
        HIP_OK(hipMalloc((void**)&gate_dec_, L_.N));
        allocs_.push_back(gate_dec_);
        HIP_OK(hipMemset(gate_dec_, 0, L_.N));
        // every wave marked until a network evaluation clears the marks
        const size_t nmix = (size_t)(L_.N + 63) / 64 * 4;
        HIP_OK(hipMalloc((void**)&blk_mixed_, nmix));
        allocs_.push_back(blk_mixed_);
        HIP_OK(hipMemset(blk_mixed_, 1, nmix));
        // the central stencils' linear operators the gated cells use
        const size_t nl = L_.n_pad() * (size_t)L_.S * L_.Q * (size_t)(L_.M - 1);
        HIP_OK(hipMalloc((void**)&lin_, nl * sizeof(double)));
        allocs_.push_back(lin_);
        dev_bytes_ += nl * sizeof(double);
        HIP_OK(dev::teno_linear_operator(teno(), info_, L_.N, (int)(L_.S * L_.Q), lin_, stream_));
    }
    HIP_OK(hipMemcpyAsync(theta_, params, kIndParams * sizeof(float), hipMemcpyHostToDevice, stream_));
    HIP_OK(dev::net_pack(theta_, net_buf_, stream_));
    synchronize();
    gate_thr_ = threshold;
    gate_ = true;
    drop_graph();  // the captured step carries the gate in its kernel arguments
    tuned_ = false;  // the gated kernel has its own best shape
}

void Solver::set_indicator_cadence(bool every_stage) {
    gate_per_stage_ = every_stage;
    drop_graph();  // the captured step carries the cadence in its kernel arguments
    tuned_ = false;  // the reuse passes' shape is tuned only for per-step cadence
}

void Solver::set_indicator_split(bool split) {
    gate_split_ = split;
    drop_graph();
    tuned_ = false;
}

uint64_t Solver::gated_cells(bool reset) {
    std::vector<uint32_t> slots(gated_slots_);
    HIP_OK(hipMemcpyAsync(slots.data(), gated_, gated_slots_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    synchronize();
    uint64_t v = 0;
    for (uint32_t s : slots) v += s;
    if (reset) {
        HIP_OK(hipMemsetAsync(gated_, 0, gated_slots_ * sizeof(uint32_t), stream_));
        synchronize();
    }
    return v;
}

// TENO decisions of the current state with the plain cut-off (the gate off),
// then the features and labels of every cell appended to the set.
void Solver::collect_features(Dataset& ds) {
    if (!L_.teno || L_.M < kIndFeatures / 4) fail("features need TENO central stencils of at least 12 cells");
    if (ds.n + L_.N > ds.cap) fail("dataset full");
    const bool g = gate_;
    gate_ = false;
    try {
        launch_teno(Un_);
        HIP_OK(dev::ind_features(geom(), teno(), Un_, dec_, ph_.gamma, ds.X, ds.y, ds.n, stream_));
        synchronize();
    } catch (...) {
        gate_ = g;
        throw;
    }
    gate_ = g;
    ds.n += L_.N;
}

void Solver::indicator_features(uint16_t* X_ref, float* y_ref) {
    Dataset tmp(L_.N);
    collect_features(tmp);
    std::vector<uint16_t> X((size_t)L_.N * kIndFeatures);
    std::vector<float> y(L_.N);
    tmp.get(0, L_.N, X.data(), y.data());
    for (uint32_t j = 0; j < L_.N; ++j) {
        const uint32_t c = L_.perm[j];
        if (X_ref) std::memcpy(X_ref + (size_t)c * kIndFeatures, &X[(size_t)j * kIndFeatures], kIndFeatures * 2);
        if (y_ref) y_ref[c] = y[j];
    }
}

void Solver::get_volumes(double* out) const {
    for (uint32_t j = 0; j < L_.N; ++j) out[L_.perm[j]] = L_.volume[j];
}

dev::CaseBufs Solver::case_bufs() const {
    dev::CaseBufs c{};
    c.g = geom();
    if (L_.teno) c.t = teno();
    c.Un = Un_;
    c.Ua = Ua_;
    c.Ub = Ub_;
    c.W = W_;
    c.prims = prims_;
    c.h = h_;
    c.fs = fs_;
    c.rhs = rhs_;
    c.lam_cell = lam_cell_;
    c.dt = dt_dev_;
    c.dt_hist = dt_hist_;
    c.lam_blk = lam_blk_;
    c.step = step_dev_;
    c.dec = dec_;
    return c;
}

int64_t Solver::device_step() {
    int64_t s0 = 0;
    HIP_OK(hipMemcpy(&s0, step_dev_, sizeof(int64_t), hipMemcpyDeviceToHost));
    return s0;
}

dev::StageArgs Solver::stage(int op, const double* Uin, double* Uout) const {
    dev::StageArgs a{};
    a.op = op;
    a.riemann = (int32_t)cfg_.riemann;
    a.fo = L_.teno ? 0 : 1;
    a.corrected = (L_.teno && L_.tspec.variant == TenoVariant::Corrected) ? 1 : 0;
    a.pp = cfg_.positivity ? 1 : 0;
    a.Uin = Uin;
    a.Un = Un_;
    a.Uout = Uout;
    a.W = W_;
    a.prims = prims_;
    a.h = h_;
    a.rhs = rhs_;
    a.fs = fs_;
    a.dt = dt_dev_;
    a.vp = vp_;
    a.vg = vg_;
    return a;
}

void Solver::mark(int tag) {
    if (!prof_on_) return;
    if (prof_used_ == prof_.size()) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        prof_.push_back({tag, e});
    }
    prof_[prof_used_].first = tag;
    HIP_OK(hipEventRecord(prof_[prof_used_].second, stream_));
    ++prof_used_;
}

void Solver::launch_teno(const double* U, bool first_stage) {
    // with the gate on, the network judges the cells on a step's first stage;
    // later stages reuse its decisions unless every stage evaluates it
    const int reuse = (gate_ && !gate_per_stage_ && !first_stage) ? 1 : 0;
    HIP_OK(dev::teno_faces(geom(), teno(), U, fs_, dec_, ph_, reuse ? tune_.teno_occ_reuse : tune_.teno_occ, reuse,
                           stream_));
    mark(kStatTenoMs);
}

void Solver::launch_flux(int op, const double* Uin, double* Uout) {
    if (cfg_.viscous) HIP_OK(dev::viscous_grad(geom(), Uin, ph_, vp_, vg_, tune_.flux_block, stream_));
    HIP_OK(dev::flux_update(geom(), stage(op, Uin, Uout), ph_, tune_.flux_block, stream_));
    mark(kStatFluxMs);
}

void Solver::stage_pass(int op, const double* Uin, double* Uout) {
    if (L_.teno) launch_teno(Uin, op == dev::kFE || op == dev::kSSP1 || op == dev::kRK4_1 || op == dev::kRhsOnly);
    launch_flux(op, Uin, Uout);
}

void Solver::dt_pass(double cfl, double fixed_dt) {
    if (fixed_dt <= 0.0)
        HIP_OK(dev::spectral_radius(geom(), Un_, prims_, ph_, lam_cell_, lam_blk_, tune_.spec_block, stream_));
    HIP_OK(dev::finalize_dt(lam_blk_, spec_blocks(), cfl, fixed_dt, dt_dev_, dt_hist_, dt_hist_cap_, step_dev_,
                            stream_));
    mark(kStatDtMs);
}

// time_integrator.cpp:57-163 as fused stage passes
void Solver::launch_step() {
    dt_pass(cfg_.cfl, cfg_.use_cfl ? 0.0 : cfg_.dt);
    switch (cfg_.integrator) {
    case Integrator::FE:
        stage_pass(dev::kFE, Un_, Ua_);
        HIP_OK(dev::copy_state(Ua_, Un_, L_.N, stream_));
        mark(kStatOtherMs);
        break;
    case Integrator::SSPRK3:
        stage_pass(dev::kSSP1, Un_, Ua_);
        stage_pass(dev::kSSP2, Ua_, Ub_);
        stage_pass(dev::kSSP3, Ub_, Un_);
        break;
    case Integrator::RK4:
        stage_pass(dev::kRK4_1, Un_, Ua_);
        stage_pass(dev::kRK4_2, Ua_, Ub_);
        stage_pass(dev::kRK4_3, Ub_, Ua_);
        stage_pass(dev::kRK4_4, Ua_, Un_);
        break;
    }
}

void Solver::record_step() {
    if (step_exec_) return;
    HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    launch_step();
    HIP_OK(hipStreamEndCapture(stream_, &step_graph_));
    HIP_OK(hipGraphInstantiate(&step_exec_, step_graph_, nullptr, nullptr, 0));
}

void Solver::synchronize() { HIP_OK(hipStreamSynchronize(stream_)); }

void Solver::set_cfl(double cfl) {
    cfg_.cfl = cfl;
    cfg_.use_cfl = true;
    drop_graph();  // the captured step carries cfl as a kernel argument
}

void Solver::device_stencils(uint32_t cell, uint32_t* out) const {
    if (!L_.teno) fail("stencils exist only for TENO");
    if (cell >= L_.N) fail("cell out of range");
    const uint32_t j = L_.iperm[cell];
    const int Mp = L_.M - 1;
    uint32_t info0 = 0;
    HIP_OK(hipMemcpy(&info0, info_ + j, 4, hipMemcpyDeviceToHost));
    for (int s = 0; s < L_.G; ++s) {
        const bool on = (info0 >> (20 + s)) & 1u;
        out[(size_t)s * L_.M] = on ? cell : 0xFFFFFFFFu;
        // member = tile list of the cell's block at the member's stored position
        const uint32_t b = j / Layout::kTileCells;
        for (int i = 0; i < Mp; ++i) {
            const size_t o = L_.mem_idx(s, i, j);
            uint32_t pos = 0;
            if (loc8_) {
                uint8_t p8 = 0;
                HIP_OK(hipMemcpy(&p8, loc8_ + o, 1, hipMemcpyDeviceToHost));
                pos = p8;
            } else {
                uint16_t p16 = 0;
                HIP_OK(hipMemcpy(&p16, loc16_ + o, 2, hipMemcpyDeviceToHost));
                pos = p16;
            }
            int32_t id = -1;
            HIP_OK(hipMemcpy(&id, blk_ids_ + L_.blk_off[b] + pos, 4, hipMemcpyDeviceToHost));
            out[(size_t)s * L_.M + 1 + i] = on ? L_.perm[id] : 0xFFFFFFFFu;
        }
    }
}

void Solver::device_matrices(uint32_t cell, double* aplus, double* ta) const {
    if (!L_.teno) fail("matrices exist only for TENO");
    if (compact_) fail("the compact tables hold A+ times the member areas; the exact tables were released");
    if (cell >= L_.N) fail("cell out of range");
    const uint32_t j = L_.iperm[cell];
    const int K = L_.K, M = L_.M, G = L_.G;
    uint32_t info0 = 0;
    HIP_OK(hipMemcpy(&info0, info_ + j, 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < (size_t)G * K * M; ++i) aplus[i] = 0.0;
    for (size_t i = 0; i < (size_t)G * M; ++i) ta[i] = 0.0;
    for (int s = 0; s < G; ++s) {
        if (!((info0 >> (20 + s)) & 1u)) continue;
        for (int i = 1; i < M; ++i) {
            HIP_OK(hipMemcpy(&ta[(size_t)s * M + i], ta_ + L_.mem_idx(s, i - 1, j), 8, hipMemcpyDeviceToHost));
            for (int k = 1; k < K; ++k)
                HIP_OK(hipMemcpy(&aplus[((size_t)s * K + k) * M + i], A_ + L_.a_idx(s, i - 1, k - 1, j), 8,
                                 hipMemcpyDeviceToHost));
        }
    }
}

void Solver::set_tune(const dev::Tune& t) {
    auto bs_ok = [](uint32_t b) { return b == 64 || b == 128 || b == 256; };
    if (!bs_ok(t.flux_block) || !bs_ok(t.spec_block)) fail("block sizes must be 64, 128 or 256 cells");
    auto occ_ok = [](int o) { return o == 0 || o == 4 || o == 5 || o == 6; };
    if (!occ_ok(t.teno_occ) || !occ_ok(t.teno_occ_reuse)) fail("TENO occupancy target must be 0, 4, 5 or 6");
    if (t.teno_face_basis != 0 && t.teno_face_basis != 1)
        fail("TENO face-basis source must be 0 (stored) or 1 (derived)");
    if (t.teno_face_basis == 1) {
        if (!cvert_) fail("derived TENO face basis: p = 2, Legendre basis, default stencil size, tiles of <= 256 cells");
        if (t.teno_occ == 6) fail("derived TENO face basis: occupancy target 0, 4 or 5");
    }
    tune_ = t;
    tuned_ = true;
    drop_graph();
}

// Times each candidate shape on the current state: launches that only write
// scratch outputs (face states, the first stage's U and accumulator, the
// radii) a few times each, keeping the fastest. The gated-cell counters are
// restored afterwards.
void Solver::autotune() {
    static std::mutex mu;
    static std::map<std::vector<int64_t>, dev::Tune> cache;
    const int fq = (int)(L_.S * L_.Q);
    const std::vector<int64_t> key = {(int64_t)L_.N, L_.K, L_.M, fq, loc8_ ? 1 : 0, gate_ ? 1 : 0,
                                      (int64_t)cfg_.riemann, L_.teno ? 1 : 0, gate_split_ ? 1 : 0,
                                      gate_per_stage_ ? 1 : 0, cfg_.viscous ? 1 : 0, compact_ ? 1 : 0};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            set_tune(it->second);
            return;
        }
    }
    drop_graph();
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    auto timed = [&](const std::function<void()>& f) {
        f();  // warm
        HIP_OK(hipEventRecord(e0, stream_));
        for (int r = 0; r < 3; ++r) f();
        HIP_OK(hipEventRecord(e1, stream_));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        return ms;
    };
    dev::Tune best = tune_;
    const bool prof = prof_on_;
    prof_on_ = false;
    // spectral radius block
    float bt = 1e30f;
    for (uint32_t bs : {64u, 128u, 256u}) {
        const float t = timed([&] {
            HIP_OK(dev::spectral_radius(geom(), Un_, prims_, ph_, lam_cell_, lam_blk_, bs, stream_));
        });
        if (t < bt) { bt = t; best.spec_block = bs; }
    }
    tune_.spec_block = best.spec_block;
    // a valid dt for the stage kernels (not a step: no history, no counter)
    HIP_OK(dev::spectral_radius(geom(), Un_, prims_, ph_, lam_cell_, lam_blk_, tune_.spec_block, stream_));
    HIP_OK(dev::finalize_dt(lam_blk_, spec_blocks(), cfg_.cfl, cfg_.use_cfl ? 0.0 : cfg_.dt, dt_dev_, nullptr, 1,
                            nullptr, stream_));
    if (L_.teno) {
        uint32_t* saved = nullptr;
        if (gate_) {
            HIP_OK(hipMalloc((void**)&saved, gated_slots_ * sizeof(uint32_t)));
            HIP_OK(hipMemcpyAsync(saved, gated_, gated_slots_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
        }
        // occupancy targets: 5 and 6 waves per SIMD for the p = 2 kernels
        // (not for split gated passes), 4 and 5 for plain p = 3 with the
        // default stencil size; the compiler's own choice (0) everywhere.
        // The plain p = 2 pass also tries its derived face basis at 0, 4 and 5.
        const bool occ_tunable = L_.K == 6 && L_.M == 12 && fq == 3 && loc8_ != nullptr && !(gate_ && gate_split_);
        const bool occ_p3 = L_.K == 10 && L_.M == 20 && fq == 6 && !gate_;
        const bool derivable = cvert_ != nullptr && !gate_ && !compact_;
        bt = 1e30f;
        for (int fb : {0, 1}) {
            if (fb == 1 && !derivable) continue;
            for (int occ : {0, 4, 5, 6}) {
                if (fb == 0 && occ != 0 && !(occ_tunable && occ >= 5) && !(occ_p3 && occ <= 5)) continue;
                if (fb == 1 && occ == 6) continue;
                tune_.teno_face_basis = fb;
                const float t = timed([&] {
                    HIP_OK(dev::teno_faces(geom(), teno(), Un_, fs_, dec_, ph_, occ, 0, stream_));
                });
                if (t < bt) { bt = t; best.teno_occ = occ; best.teno_face_basis = fb; }
            }
        }
        tune_.teno_occ = best.teno_occ;
        tune_.teno_face_basis = best.teno_face_basis;
        HIP_OK(dev::teno_faces(geom(), teno(), Un_, fs_, dec_, ph_, tune_.teno_occ, 0, stream_));
        if (gate_ && !gate_per_stage_) {  // passes that reuse the decisions just recorded
            bt = 1e30f;
            for (int occ : {0, 5, 6}) {
                if (occ != 0 && !occ_tunable) continue;
                const float t = timed([&] {
                    HIP_OK(dev::teno_faces(geom(), teno(), Un_, fs_, dec_, ph_, occ, 1, stream_));
                });
                if (t < bt) { bt = t; best.teno_occ_reuse = occ; }
            }
        }
        if (saved) {
            HIP_OK(hipMemcpyAsync(gated_, saved, gated_slots_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
            synchronize();
            HIP_OK(hipFree(saved));
        }
    }
    // a first stage (never the final one, which would rewrite the primitives
    // the next dt is computed from); FE's only stage is timed as the residual
    const int op = cfg_.integrator == Integrator::FE ? dev::kRhsOnly
                   : cfg_.integrator == Integrator::RK4 ? dev::kRK4_1 : dev::kSSP1;
    bt = 1e30f;
    for (uint32_t bs : {64u, 128u, 256u}) {
        const float t = timed([&] {
            if (cfg_.viscous) HIP_OK(dev::viscous_grad(geom(), Un_, ph_, vp_, vg_, bs, stream_));
            HIP_OK(dev::flux_update(geom(), stage(op, Un_, Ua_), ph_, bs, stream_));
        });
        if (t < bt) { bt = t; best.flux_block = bs; }
    }
    prof_on_ = prof;
    HIP_OK(hipEventDestroy(e0));
    HIP_OK(hipEventDestroy(e1));
    synchronize();
    set_tune(best);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = best;
}

void check_dt_config(const Config& c) {
    if (c.use_cfl && !(c.cfl >= 0.0)) fail("dt negative: " + std::to_string(c.cfl) + ".");
    if (!c.use_cfl && c.dt < 0.0) fail("dt negative: " + std::to_string(c.dt) + ".");
    if (!c.use_cfl && !(c.dt > 0.0)) fail("a fixed dt must be positive (or give a CFL number)");
}

void Solver::check_dt_mode() const { check_dt_config(cfg_); }

int64_t Solver::stream_step() {
    int64_t s0 = 0;
    HIP_OK(hipMemcpyAsync(&s0, step_dev_, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    synchronize();
    return s0;
}

// The device writes each step's dt into a ring of dt_hist_cap_ entries, so a
// run is issued in chunks of at most that many steps, each read back before
// the next chunk overwrites the ring.
void Solver::run_steps(int64_t n, double* dts_out) {
    if (n <= 0) return;
    check_dt_mode();
    ensure_tuned();
    if (use_graphs_) record_step();
    for (int64_t done = 0; done < n;) {
        const int64_t m = std::min<int64_t>(n - done, dt_hist_cap_);
        const int64_t s0 = stream_step();
        for (int64_t i = 0; i < m; ++i) {
            if (use_graphs_) HIP_OK(hipGraphLaunch(step_exec_, stream_));
            else launch_step();
        }
        synchronize();
        finish_steps(s0, m, dts_out ? dts_out + done : nullptr);
        done += m;
    }
}

void Solver::run_steps_profiled(int64_t n, double* stats) {
    for (int k = 0; k < kStatCount; ++k) stats[k] = 0.0;
    if (n <= 0) return;
    check_dt_mode();
    ensure_tuned();
    for (int64_t done = 0; done < n;) {
        const int64_t m = std::min<int64_t>(n - done, dt_hist_cap_);
        const int64_t s0 = stream_step();
        prof_on_ = true;
        prof_used_ = 0;
        try {
            mark(-1);
            for (int64_t i = 0; i < m; ++i) launch_step();
        } catch (...) {
            prof_on_ = false;
            throw;
        }
        prof_on_ = false;
        synchronize();
        // launch i ran between the events recorded before and after it
        for (size_t i = 1; i < prof_used_; ++i) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, prof_[i - 1].second, prof_[i].second));
            const int tag = prof_[i].first;
            stats[tag] += ms;
            stats[tag + 1] += 1.0;
        }
        float tot = 0.0f;
        HIP_OK(hipEventElapsedTime(&tot, prof_[0].second, prof_[prof_used_ - 1].second));
        stats[kStatTotalMs] += tot;
        finish_steps(s0, m, nullptr);
        done += m;
    }
}

// The n <= dt_hist_cap_ dt values of the steps s0 .. s0 + n - 1, which the
// device wrote at ring positions (s0 + i) % dt_hist_cap_.
void Solver::finish_steps(int64_t s0, int64_t n, double* dts_out) {
    if (n > dt_hist_cap_) fail("internal: more steps than the dt history holds");
    std::vector<double> d((size_t)n);
    const int64_t first = s0 % dt_hist_cap_;
    const int64_t part = std::min<int64_t>(n, dt_hist_cap_ - first);
    HIP_OK(hipMemcpy(d.data(), dt_hist_ + first, sizeof(double) * part, hipMemcpyDeviceToHost));
    if (n > part) HIP_OK(hipMemcpy(d.data() + part, dt_hist_, sizeof(double) * (n - part), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) {
        const double dt = d[(size_t)i];
        if (dts_out) dts_out[i] = dt;
        if (dt < 0.0) fail("dt negative: " + std::to_string(dt) + ".");
        t_ += dt;
        dt_ = dt;
    }
    step_ += n;
}

void Solver::set_dt_history(int64_t cap) {
    if (cap < 1 || cap > kDtHistCap) fail("the dt history holds 1 to 1048576 steps");
    synchronize();
    drop_graph();  // the captured finalize kernel carries the ring's address and length
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, sizeof(double) * cap));
    HIP_OK(hipMemset(p, 0, sizeof(double) * cap));
    for (void*& a : allocs_)
        if (a == dt_hist_) {
            HIP_OK(hipFree(a));
            a = p;
        }
    dev_bytes_ += sizeof(double) * (uint64_t)cap;
    dev_bytes_ -= sizeof(double) * (uint64_t)dt_hist_cap_;
    dt_hist_ = static_cast<double*>(p);
    dt_hist_cap_ = cap;
}

double Solver::calc_dt() {
    check_dt_mode();
    if (cfg_.use_cfl) {
        HIP_OK(dev::spectral_radius(geom(), Un_, prims_, ph_, lam_cell_, lam_blk_, tune_.spec_block, stream_));
        HIP_OK(dev::finalize_dt(lam_blk_, spec_blocks(), cfg_.cfl, 0.0, dt_dev_, nullptr, 1, nullptr, stream_));
    } else {
        HIP_OK(dev::finalize_dt(lam_blk_, spec_blocks(), 0.0, cfg_.dt, dt_dev_, nullptr, 1, nullptr, stream_));
    }
    HIP_OK(hipMemcpyAsync(&dt_, dt_dev_, sizeof(double), hipMemcpyDeviceToHost, stream_));
    synchronize();
    if (dt_ < 0.0) fail("dt negative: " + std::to_string(dt_) + ".");
    return dt_;
}

void Solver::take_step() {
    switch (cfg_.integrator) {
    case Integrator::FE:
        stage_pass(dev::kFE, Un_, Ua_);
        HIP_OK(dev::copy_state(Ua_, Un_, L_.N, stream_));
        mark(kStatOtherMs);
        break;
    case Integrator::SSPRK3:
        stage_pass(dev::kSSP1, Un_, Ua_);
        stage_pass(dev::kSSP2, Ua_, Ub_);
        stage_pass(dev::kSSP3, Ub_, Un_);
        break;
    case Integrator::RK4:
        stage_pass(dev::kRK4_1, Un_, Ua_);
        stage_pass(dev::kRK4_2, Ua_, Ub_);
        stage_pass(dev::kRK4_3, Ub_, Ua_);
        stage_pass(dev::kRK4_4, Ua_, Un_);
        break;
    }
    synchronize();
    step_++;
    t_ += dt_;
}

void Solver::calc_rhs(double* rhs_ref) {
    if (L_.teno) launch_teno(Un_);
    launch_flux(dev::kRhsOnly, Un_, nullptr);
    std::vector<double> r(4 * (size_t)L_.N);
    HIP_OK(hipMemcpyAsync(r.data(), rhs_, r.size() * 8, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t j = 0; j < L_.N; ++j) std::memcpy(&rhs_ref[4 * (size_t)L_.perm[j]], &r[4 * (size_t)j], 32);
}

void Solver::set_state(const double* U, const double* prim5) {
    const size_t N = L_.N;
    std::vector<double> u(4 * N), w(4 * N), h(N);
    for (uint32_t j = 0; j < N; ++j) {
        const uint32_t c = L_.perm[j];
        std::memcpy(&u[4 * (size_t)j], &U[4 * (size_t)c], 32);
        double W5[5];
        if (prim5) std::memcpy(W5, &prim5[5 * (size_t)c], 40);
        else host_prims(ph_, &U[4 * (size_t)c], W5);
        std::memcpy(&w[4 * (size_t)j], W5, 32);
        h[j] = W5[4];
    }
    HIP_OK(hipMemcpy(Un_, u.data(), u.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(prims_, w.data(), w.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(h_, h.data(), h.size() * 8, hipMemcpyHostToDevice));
}

void Solver::get_state(double* U, double* prim5) {
    const size_t N = L_.N;
    std::vector<double> u(4 * N), w(4 * N), h(N);
    synchronize();
    HIP_OK(hipMemcpy(u.data(), Un_, u.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(w.data(), prims_, w.size() * 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(h.data(), h_, h.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < N; ++j) {
        const uint32_t c = L_.perm[j];
        if (U) std::memcpy(&U[4 * (size_t)c], &u[4 * (size_t)j], 32);
        if (prim5) {
            std::memcpy(&prim5[5 * (size_t)c], &w[4 * (size_t)j], 32);
            prim5[5 * (size_t)c + 4] = h[j];
        }
    }
}

// solver_initialize.cpp:35-186
void Solver::init_solution() {
    const uint32_t N = mesh_.n_cells;
    std::vector<double> U(4 * (size_t)N), W(5 * (size_t)N);
    auto put = [&](uint32_t c, double rho, double u0, double u1, double p, double T) {
        const double e = ph_.cv * T;
        const double E = e + 0.5 * (u0 * u0 + u1 * u1);
        const double h = e + p / rho;
        double* Uc = &U[4 * (size_t)c];
        Uc[0] = rho; Uc[1] = rho * u0; Uc[2] = rho * u1; Uc[3] = rho * E;
        double* Wc = &W[5 * (size_t)c];
        Wc[0] = u0; Wc[1] = u1; Wc[2] = p; Wc[3] = T; Wc[4] = h;
    };
    const InitSpec& I = cfg_.init;
    if (I.kind == InitSpec::Constant) {
        const double rho = I.p / (ph_.R * I.T);
        for (uint32_t c = 0; c < N; ++c) put(c, rho, I.u[0], I.u[1], I.p, I.T);
    } else if (I.kind == InitSpec::Quadrant) {
        for (uint32_t c = 0; c < N; ++c) {
            const bool r = mesh_.cc(c, 0) >= I.xs, t = mesh_.cc(c, 1) >= I.ys;
            const int q = r ? (t ? 0 : 3) : (t ? 1 : 2);
            const double* s = &I.states[4 * q];
            const double rho = s[0], p = s[3];
            put(c, rho, s[1], s[2], p, p / (rho * ph_.R));
        }
    } else {
        Expression ux(I.u_expr[0]), uy(I.u_expr[1]);
        std::unique_ptr<Expression> er, ep, eT;
        if (!I.rho_expr.empty()) er.reset(new Expression(I.rho_expr));
        if (!I.p_expr.empty()) ep.reset(new Expression(I.p_expr));
        if (!I.T_expr.empty()) eT.reset(new Expression(I.T_expr));
        for (uint32_t c = 0; c < N; ++c) {
            const double x = mesh_.cc(c, 0), y = mesh_.cc(c, 1);
            const double u0 = ux.eval(x, y), u1 = uy.eval(x, y);
            double rho = 0.0, p = 0.0, T = 0.0;
            if (er) rho = er->eval(x, y);
            if (ep) p = ep->eval(x, y);
            if (eT) T = eT->eval(x, y);
            if (er && ep) T = p / (rho * ph_.R);
            else if (er && eT) p = rho * ph_.R * T;
            else if (ep && eT) rho = p / (ph_.R * T);
            else fail("Exactly two of rho, p, and T must be specified for initialization: analytical.");
            put(c, rho, u0, u1, p, T);
        }
    }
    set_state(U.data(), W.data());
}

void Solver::get_face_values(double* out) { face_values_from(Un_, out); }

void Solver::upload_state(const double* U_ref, double* dst_dev) {
    if (!U_ref) fail("state is NULL");
    const size_t N = L_.N;
    std::vector<double> u(4 * N);
    for (uint32_t j = 0; j < N; ++j) std::memcpy(&u[4 * (size_t)j], &U_ref[4 * (size_t)L_.perm[j]], 32);
    synchronize();
    HIP_OK(hipMemcpy(dst_dev, u.data(), u.size() * 8, hipMemcpyHostToDevice));
}

// The stage buffer Ua_ carries the caller's state: between steps it holds
// nothing a later step reads before writing it.
void Solver::calc_face_values_of(const double* U_ref, double* out) {
    if (!out) fail("face_solution is NULL");
    upload_state(U_ref, Ua_);
    if (L_.teno) launch_teno(Ua_);
    face_values_from(Ua_, out);
}

void Solver::calc_rhs_device(const double* U_dev, int parts, double* rhs_dev) {
    if (parts < 1 || parts > 3) fail("calc_rhs: parts must be 1 (interior), 2 (boundaries) or 3 (both)");
    if (!U_dev || !rhs_dev) fail("calc_rhs: NULL device array");
    if (L_.teno) launch_teno(U_dev);
    if (cfg_.viscous) HIP_OK(dev::viscous_grad(geom(), U_dev, ph_, vp_, vg_, tune_.flux_block, stream_));
    dev::StageArgs a = stage(dev::kRhsOnly, U_dev, nullptr);
    a.rhs = rhs_dev;
    HIP_OK(dev::rhs_parts(geom(), a, ph_, parts, tune_.flux_block, stream_));
    mark(kStatFluxMs);
}

void Solver::calc_rhs_of(const double* U_ref, int parts, double* rhs_ref) {
    if (!rhs_ref) fail("rhs is NULL");
    upload_state(U_ref, Ua_);
    calc_rhs_device(Ua_, parts, rhs_);
    std::vector<double> r(4 * (size_t)L_.N);
    HIP_OK(hipMemcpyAsync(r.data(), rhs_, r.size() * 8, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t j = 0; j < L_.N; ++j) std::memcpy(&rhs_ref[4 * (size_t)L_.perm[j]], &r[4 * (size_t)j], 32);
}

void Solver::face_values_from(const double* cells_dev, double* out) {
    const uint32_t Q = L_.Q, N = L_.N;
    const size_t nf = mesh_.n_faces;
    std::fill(out, out + nf * Q * 2 * 4, NAN);
    std::vector<double> src;
    synchronize();
    if (L_.teno) {
        src.resize(4 * (size_t)L_.S * Q * N);
        HIP_OK(hipMemcpy(src.data(), fs_, src.size() * 8, hipMemcpyDeviceToHost));
    } else {
        src.resize(4 * (size_t)N);
        HIP_OK(hipMemcpy(src.data(), cells_dev, src.size() * 8, hipMemcpyDeviceToHost));
    }
    for (uint32_t f = 0; f < mesh_.n_faces_real; ++f)
        for (int side = 0; side < 2; ++side) {
            const int32_t c = mesh_.cof(f, side);
            if (c < 0) continue;
            const uint32_t j = L_.iperm[c];
            uint32_t l = 0;
            for (uint32_t k = 0; k < L_.S; ++k)
                if (mesh_.face_of_cell(c, k) == f) l = k;
            for (uint32_t q = 0; q < Q; ++q) {
                const double* s = L_.teno ? &src[4 * ((size_t)(l * Q + q) * N + j)] : &src[4 * (size_t)j];
                std::memcpy(&out[((f * (size_t)Q + q) * 2 + side) * 4], s, 32);
            }
        }
}

void Solver::get_gate_decisions(uint8_t* out) {
    if (!gate_dec_) fail("the learned indicator has not been set");
    std::vector<uint8_t> d(L_.N);
    synchronize();
    HIP_OK(hipMemcpy(d.data(), gate_dec_, d.size(), hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < L_.N; ++j) out[L_.perm[j]] = d[j];
}

void Solver::get_decisions(uint8_t* out) {
    std::vector<uint8_t> d(L_.N);
    synchronize();
    HIP_OK(hipMemcpy(d.data(), dec_, d.size(), hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < L_.N; ++j) out[L_.perm[j]] = d[j];
}

// solver.cpp:580-590, :705: cfl_local = dt * lambda, filled only when the
// step comes from the CFL number; with a fixed dt the reference never runs
// the spectral radius and its (zero-initialised) cfl_local stays zero
void Solver::get_cfl_local(double* out) {
    if (!cfg_.use_cfl) {
        std::fill(out, out + L_.N, 0.0);
        return;
    }
    std::vector<double> l(L_.N);
    synchronize();
    HIP_OK(hipMemcpy(l.data(), lam_cell_, l.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < L_.N; ++j) out[L_.perm[j]] = l[j] * dt_;  // KokkosBlas::scal
}

void Solver::check_fields() {
    if (!cfg_.check_nan) return;
    const uint32_t N = mesh_.n_cells;
    std::vector<double> U(4 * (size_t)N), W(5 * (size_t)N);
    get_state(U.data(), W.data());
    for (uint32_t c = 0; c < N; ++c) {
        bool nan = false;
        for (int i = 0; i < 4; ++i) nan |= std::isnan(U[4 * (size_t)c + i]);
        for (int i = 0; i < 5; ++i) nan |= std::isnan(W[5 * (size_t)c + i]);
        if (nan) {
            std::ostringstream msg;
            msg << "NaN found in solution.\nt: " << t_ << "\nstep: " << step_ << "\ni_cell: " << c
                << "\n> x: " << mesh_.cc(c, 0) << "\n> y: " << mesh_.cc(c, 1) << "\n";
            fail(msg.str());
        }
    }
}

// solver.cpp:352-403 (checks and VTU output are not part of this engine)
// solver.cpp:421-425: the step-info print every check_interval steps
void Solver::print_step_info() const {
    std::printf("Starting step: %lld time: %g dt: %g\n", (long long)step_, t_, dt_);
}

// solver.cpp:500-504 and data_writer.cpp:78-92: every writer whose interval
// divides the step (or all of them when forced) writes the current state
void Solver::write_data(bool force) {
    bool any = false;
    for (const WriteSpec& w : cfg_.writers) any |= force || step_ % w.interval == 0;
    if (!any) return;
    const size_t N = mesh_.n_cells;
    std::vector<double> U(4 * N), W(5 * N), cfl(N);
    get_state(U.data(), W.data());
    get_cfl_local(cfl.data());
    for (const WriteSpec& w : cfg_.writers)
        if (force || step_ % w.interval == 0) write_vtu(vtu_name(w.prefix, (uint64_t)step_), mesh_, w.variables,
                                                      U.data(), W.data(), cfl.data());
}

// solver.cpp:352-403. With no time or wall-clock stop, no NaN check and no
// output, the steps run as one graph-replayed batch; otherwise step by step
// (each step still one graph launch) so the stop conditions, checks and
// writers see every step.
int Solver::run() {
    auto done = [&]() {
        bool d = false;
        if (cfg_.n_steps > 0) d |= step_ >= cfg_.n_steps;
        if (cfg_.t_stop > 0) d |= t_ >= cfg_.t_stop;
        if (cfg_.t_wall_stop > 0) d |= (wall_seconds() - wall0_) >= cfg_.t_wall_stop;
        return d;
    };
    const bool batch = cfg_.t_stop <= 0 && cfg_.t_wall_stop <= 0 && !cfg_.check_nan && cfg_.writers.empty();
    if (batch) {
        const int64_t n = cfg_.n_steps - step_;
        if (n > 0) run_steps(n, nullptr);
        return 0;
    }
    write_data(true);
    while (!done()) {
        if (cfg_.check_interval > 0 && step_ % cfg_.check_interval == 0) print_step_info();
        run_steps(1, nullptr);
        check_fields();
        write_data(false);
    }
    write_data(true);
    return 0;
}

}  // namespace mlr
