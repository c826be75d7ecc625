// Batched ensembles: many independent cases (own meshes, states and CFL
// time steps) advanced together, every kernel phase of a step one launch over
// all cases' blocks driven by a device table of per-case descriptors and
// block offsets, and the whole step captured in one hipGraph. Each case's
// arithmetic is exactly its standalone solver's, so results are bit-identical
// to running the cases one by one.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "solver.hpp"

namespace mlr {

struct MeshSpec {
    MeshType type = MeshType::CartesianTri;
    uint32_t nx = 64, ny = 64;
    double Lx = 1.0, Ly = 1.0, jitter = 0.0;
    uint64_t seed = 0;
};

class Ensemble {
public:
    // Builds one solver per mesh with the shared numerics of `base`.
    Ensemble(const Config& base, const std::vector<MeshSpec>& meshes);
    ~Ensemble();
    Ensemble(const Ensemble&) = delete;
    Ensemble& operator=(const Ensemble&) = delete;

    int size() const { return (int)cases_.size(); }
    Solver& solver(int k) { return *cases_.at((size_t)k); }
    uint64_t n_cells() const;

    // the learned indicator on every case (params NULL: off)
    void set_indicator(const float* params, double threshold);
    // as Solver::set_indicator_cadence / set_indicator_split
    void set_indicator_cadence(bool every_stage);
    void set_indicator_split(bool split);
    // launch shapes (flux and spectral blocks share one size here), timed on
    // the cases' current states before the first steps unless pinned
    void autotune();
    void set_tune(const dev::Tune& t);
    const dev::Tune& tune() const { return tune_; }
    // n steps of every case, graph-replayed
    void run_steps(int64_t n);
    // the same launches one at a time with an event after each (Solver's StepStats layout)
    void run_steps_profiled(int64_t n, double* stats);
    // run_steps as one persistent cooperative launch per step (every phase of
    // every case, grid barriers between phases) instead of the graph of
    // per-phase launches: plain TENO at p = 2 with the default stencil size
    // and tile width, or first order. Pins flux and spectral blocks to 256
    // cells. Profiled runs then take each phase's time from device clock
    // stamps at the barriers (the last stage's flux phase gains a barrier).
    void set_persistent(bool on);
    bool persistent() const { return persistent_; }
    int persistent_grid() const { return persistent_grid_; }
    // longest wait of a block at a grid barrier of the persistent step before
    // the step is abandoned and run_steps fails (default 2000 ms)
    void set_persistent_wait_limit(double ms);

private:
    void upload_table();
    void launch_step(bool prof);
    void launch_persistent(long long* clocks);
    void run_steps_profiled_persistent(int64_t n, double* stats);
    void mark(int tag);
    void drop_graph();
    void finish(int64_t s0, int64_t n);

    Config cfg_;
    std::vector<std::unique_ptr<Solver>> cases_;
    dev::CaseBufs* d_cases_ = nullptr;
    uint32_t teno_blocks_ = 0, flux_blocks_ = 0;
    int K_ = 0, M_ = 0, FQ_ = 0;
    bool small_ = true, teno_ = false, gated_ = false, gate_per_stage_ = false, split_ = false;
    bool derivable_ = false;  // every case holds its cells' node coordinates (derived face basis)
    dev::Physics ph_{};
    dev::Tune tune_{};
    bool tuned_ = false;
    hipStream_t stream_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    bool persistent_ = false;
    int persistent_grid_ = 0;
    dev::StepBarrier* d_bar_ = nullptr;
    double wait_limit_ms_ = 2000.0;
    long long spin_ticks_ = 0;
    void check_barrier();
    bool prof_on_ = false;
    std::vector<std::pair<int, hipEvent_t>> prof_;
    size_t prof_used_ = 0;
};

}  // namespace mlr
