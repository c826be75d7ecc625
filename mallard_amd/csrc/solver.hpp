// Host solver mirroring the reference's Solver (solver.h:30-279): TOML
// input, mesh / physics / numerics / boundaries / run-parameter init, the run
// loop, and the residual, dt and primitive updates, all executed by the HIP
// kernels on one GPU through one stream.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "io.hpp"
#include "layout.hpp"
#include "mesh.hpp"
#include "toml.hpp"

namespace mlr {

struct Dataset;

// Layout of the stats array filled by Solver::run_steps_profiled.
enum StepStats {
    kStatTotalMs = 0,   // first launch start to last launch end, event time
    kStatTenoMs, kStatTenoN,      // TENO face-state kernels: summed ms, launches
    kStatFluxMs, kStatFluxN,      // flux + stage-update kernels
    kStatDtMs, kStatDtN,          // spectral radius + dt finalisation
    kStatOtherMs, kStatOtherN,    // copies (forward Euler)
    kStatCount
};

struct InitSpec {
    enum Kind { Constant, Analytical, Quadrant } kind = Constant;
    double u[2] = {0.0, 0.0};
    double p = 0.0, T = 0.0;                          // constant
    std::string u_expr[2], rho_expr, p_expr, T_expr;  // analytical
    double xs = 0.5, ys = 0.5, states[16] = {};       // quadrant (rho,u,v,p) UR, UL, LL, LR
};

struct Config {
    // [mesh]
    MeshType mesh_type = MeshType::File;
    uint32_t nx = 100, ny = 100;
    double Lx = 1.0, Ly = 1.0, jitter = 0.0;
    uint64_t seed = 0;
    // [physics]
    double gamma = 1.4, p_ref = 101325.0, T_ref = 298.15, rho_ref = 1.225, p_min = -1e20, p_max = 1e20;
    // Navier-Stokes (no reference counterpart): Sutherland viscosity
    // mu = mu_ref (T / T_mu)^1.5 (T_mu + S) / (T + S) (constant mu_ref when S < 0),
    // heat conductivity mu cp / Pr
    bool viscous = false;
    double mu_ref = 0.0, T_mu = 1.0, S_mu = -1.0, Pr = 0.72;
    // [numerics]
    Recon recon = Recon::FO;
    TenoSpec teno;
    Riemann riemann = Riemann::HLLC;
    Integrator integrator = Integrator::SSPRK3;
    // positivity-preserving flux limiter (Hu, Adams & Shu 2013; no reference
    // counterpart): forward Euler and SSP-RK3, Euler equations
    bool positivity = false;
    bool check_nan = false;
    std::vector<BoundarySpec> bcs;
    // [run]
    bool use_cfl = true;
    double cfl = 0.5, dt = 0.0;
    int64_t n_steps = -1;
    double t_stop = -1.0, t_wall_stop = -1.0;
    // [output], [[write_data]]
    uint32_t check_interval = 1;
    std::vector<WriteSpec> writers;
    InitSpec init;
    int n_threads = 0;  // host precompute threads (0: OpenMP default)
};

// Parses a Mallard TOML input with the reference's keys and defaults.
Config config_from_toml(const toml::Value& in);

// The time-step mode of a configuration, checked before any step
// (solver.cpp:580-590): a negative CFL number or fixed dt is refused with the
// reference's "dt negative" message, and a fixed dt of zero too (the device's
// dt pass reads a zero fixed dt as "use the CFL number"). Solvers and
// ensembles share it.
void check_dt_config(const Config& c);

class Ensemble;

class Solver {
    friend class Ensemble;

public:
    Solver();
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // solver.cpp:39-80
    int init(const std::string& input_file);
    void init(const Config& cfg);                 // generated mesh from cfg
    void init(const Config& cfg, Mesh&& mesh);    // caller mesh (reference numbering)
    // solver.cpp:352-373 (no VTU output)
    int run();

    // reference-numbered host state I/O
    void set_state(const double* U, const double* prim5);
    void get_state(double* U, double* prim5);
    void init_solution();
    void calc_rhs(double* rhs_ref);                 // R(Un) into host rhs (reference order)
    double calc_dt();                               // solver.cpp:580-590
    void take_step();                               // one step with the current dt mode
    // n steps, dt from CFL (or fixed dt) computed on the device; the step is
    // captured once into a hipGraph and replayed
    void run_steps(int64_t n, double* dts_out);
    // the same n steps launched one kernel at a time with a HIP event recorded
    // on the stream after each launch; stats receives the event-measured totals
    // (see StepStats)
    void run_steps_profiled(int64_t n, double* stats);
    // length of the device ring of per-step dt values (default 2^20 steps):
    // runs are issued in chunks of at most this many steps
    void set_dt_history(int64_t cap);
    void get_face_values(double* out);              // [n_faces alloc][Q][2][4], reference numbering
    // FaceReconstruction::calc_face_values(solution, face_solution) on a
    // caller's state (host, reference numbering) into out as get_face_values;
    // the solver's state, time and dt are untouched
    void calc_face_values_of(const double* U_ref, double* out);
    // Solver::calc_rhs on a caller's state (host, reference numbering); parts:
    // 1 interior faces (calc_rhs_interior), 2 boundary faces
    // (calc_rhs_boundaries), 3 both, each divided by the volume
    void calc_rhs_of(const double* U_ref, int parts, double* rhs_ref);
    // the same on device arrays [N][4] in this solver's numbering, queued on
    // its stream (no host copies, no synchronisation)
    void calc_rhs_device(const double* U_dev, int parts, double* rhs_dev);
    void get_decisions(uint8_t* out);               // [N] bit v = central stencil for variable v
    void get_cfl_local(double* out);
    void check_fields();                            // solver.cpp:464-498
    void write_data(bool force);                    // solver.cpp:500-504 (VTU writers)
    void print_step_info() const;
    void synchronize();
    void set_cfl(double cfl);
    // stencil ids of one reference cell as uploaded (reference numbering, G x M)
    void device_stencils(uint32_t cell, uint32_t* ids) const;
    // uploaded A+ (G x K x M, zero first row / column restored) and member
    // areas (G x M, the target's area first) of one reference cell
    void device_matrices(uint32_t cell, double* aplus, double* ta) const;

    const Config& config() const { return cfg_; }
    const Mesh& mesh() const { return mesh_; }
    const Layout& layout() const { return L_; }
    double time() const { return t_; }
    int64_t step() const { return step_; }
    double dt() const { return dt_; }
    hipStream_t stream() const { return stream_; }
    void set_graphs(bool on) { use_graphs_ = on; }

    // kernel-level entry points used by the benchmark's timing harness
    void launch_step();                 // one full step, asynchronously
    void launch_teno(const double* U, bool first_stage = true);  // one TENO face-state pass
    void launch_flux(int op, const double* Uin, double* Uout);
    double* d_state() { return Un_; }
    uint64_t device_bytes() const { return dev_bytes_; }  // all device allocations

    // learned troubled-cell indicator: FP32 master parameters (7361) and the
    // smooth-probability threshold; NULL turns the gate off
    void set_indicator(const float* params, double threshold);
    bool indicator_on() const { return gate_; }
    // false (default): the network judges the cells once per step, on its
    // first stage, and the later stages reuse the decisions; true: every stage
    void set_indicator_cadence(bool every_stage);
    // true (default): gated passes run split, the gated cells in one kernel
    // and the ungated ones in another (identical results)
    void set_indicator_split(bool split);
    void get_gate_decisions(uint8_t* out);  // last network evaluation, reference numbering
    uint64_t gated_cells(bool reset);      // cells judged smooth since the last reset
    void collect_features(Dataset& ds);    // features + TENO labels of the current state
    void indicator_features(uint16_t* X_ref, float* y_ref);  // same, reference numbering
    void get_volumes(double* out) const;   // cell volumes, reference numbering
    int tile_pos_bytes() const { return loc8_ ? 1 : 2; }
    double* d_fs() { return fs_; }

    // compact TENO tables (corrected weighting, p = 2): the member areas
    // folded into A+ and the face-point basis values from the reference
    // triangle's edges; one way (the exact tables are released)
    void compact_tables();
    bool compact() const { return compact_; }

    // Launch shapes (results are identical for every choice). autotune()
    // times the candidates on the current state and keeps the fastest,
    // cached per mesh size and numerics for the process; it runs by itself
    // before the first steps unless set_tune() pinned a choice.
    void autotune();
    void set_tune(const dev::Tune& t);
    const dev::Tune& tune() const { return tune_; }

private:
    void setup(Mesh&& mesh);
    void upload();
    void precompute_aplus();
    void free_device();
    void record_step();
    dev::Geom geom() const;
    dev::Teno teno() const;
    dev::StageArgs stage(int op, const double* Uin, double* Uout) const;
    void stage_pass(int op, const double* Uin, double* Uout);
    void dt_pass(double cfl, double fixed_dt);
    void check_dt_mode() const;
    void mark(int tag);
    void drop_graph();
    dev::CaseBufs case_bufs() const;   // this solver as one case of an ensemble
    int64_t device_step();
    int64_t stream_step();  // the device step counter, read on the context's stream
    void finish_steps(int64_t s0, int64_t n, double* dts_out);
    void face_values_from(const double* cells_dev, double* out);  // fs_ (TENO) or the cell states (FO)
    void upload_state(const double* U_ref, double* dst_dev);     // reference numbering -> dst_dev

    Config cfg_;
    Mesh mesh_;
    Layout L_;
    dev::Physics ph_{};
    hipStream_t stream_ = nullptr;
    bool use_graphs_ = true;

    // device memory
    std::vector<void*> allocs_;
    uint64_t dev_bytes_ = 0;
    double *volume_ = nullptr, *nx_ = nullptr, *ny_ = nullptr, *area_ = nullptr, *dxn_ = nullptr, *qw_ = nullptr;
    // viscous runs: slot vectors, (u, v, T, mu) and cell gradients of the stage input
    double *dcx_ = nullptr, *dcy_ = nullptr, *mfx_ = nullptr, *mfy_ = nullptr, *vp_ = nullptr, *vg_ = nullptr;
    int32_t* nb_ = nullptr;
    uint32_t* info_ = nullptr;
    dev::BCParam* bc_ = nullptr;
    uint8_t* loc8_ = nullptr;
    uint16_t* loc16_ = nullptr;
    uint32_t* blk_off_ = nullptr;
    int32_t* blk_ids_ = nullptr;
    double *ta_ = nullptr, *A_ = nullptr, *fbasis_ = nullptr;
    bool compact_ = false;
    double* fbc_ = nullptr;  // compact tables: basis values per reference edge
    // derived face basis: [N][6] node coordinates of each cell (null where
    // the plain TENO pass cannot derive it)
    double* cvert_ = nullptr;
    double *Un_ = nullptr, *Ua_ = nullptr, *Ub_ = nullptr, *W_ = nullptr, *prims_ = nullptr, *h_ = nullptr;
    double *fs_ = nullptr, *rhs_ = nullptr, *lam_cell_ = nullptr, *dt_dev_ = nullptr, *dt_hist_ = nullptr;
    unsigned long long* lam_blk_ = nullptr;  // per-block spectral maxima (room for 64-cell blocks)
    uint32_t spec_blocks() const { return (L_.N + tune_.spec_block - 1) / tune_.spec_block; }
    dev::Tune tune_{};
    bool tuned_ = false;
    void ensure_tuned() {
        if (!tuned_) autotune();
    }
    int64_t* step_dev_ = nullptr;
    uint32_t* gated_ = nullptr;  // per-wave-slot smooth counts of the gated TENO kernel
    uint8_t* gate_dec_ = nullptr;  // per-cell decisions of the last network evaluation
    double* lin_ = nullptr;        // central stencils' linear operators (gated cells)
    bool gate_per_stage_ = false;
    bool gate_split_ = true;
    uint8_t* blk_mixed_ = nullptr;  // split passes: per wave slot, 1 when it holds an ungated cell
    size_t gated_slots_ = 0;
    void* net_buf_ = nullptr;
    float* theta_ = nullptr;
    bool gate_ = false;
    double gate_thr_ = 0.9;
    uint8_t* dec_ = nullptr;
    int64_t dt_hist_cap_ = 0;

    // launch profiling: (kernel tag, event recorded after that launch)
    bool prof_on_ = false;
    std::vector<std::pair<int, hipEvent_t>> prof_;
    size_t prof_used_ = 0;

    hipGraphExec_t step_exec_ = nullptr;
    hipGraph_t step_graph_ = nullptr;

    double t_ = 0.0, dt_ = 0.0;
    int64_t step_ = 0;
    double wall0_ = 0.0;
};

}  // namespace mlr
