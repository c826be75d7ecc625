#include "ensemble.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <string>

namespace mlr {

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

Ensemble::Ensemble(const Config& base, const std::vector<MeshSpec>& meshes) : cfg_(base) {
    if (meshes.empty()) fail("an ensemble needs at least one case");
    if (cfg_.integrator == Integrator::FE) fail("ensembles run the RK4 and SSPRK3 integrators");
    if (cfg_.viscous) fail("ensembles run the Euler equations");
    check_dt_config(cfg_);  // before any case allocates
    cases_.reserve(meshes.size());
    for (const MeshSpec& ms : meshes) {
        Config c = cfg_;
        c.mesh_type = ms.type;
        c.nx = ms.nx;
        c.ny = ms.ny;
        c.Lx = ms.Lx;
        c.Ly = ms.Ly;
        c.jitter = ms.jitter;
        c.seed = ms.seed;
        cases_.emplace_back(new Solver());
        cases_.back()->init(c);
    }
    const Solver& s0 = *cases_[0];
    teno_ = s0.L_.teno;
    ph_ = s0.ph_;
    if (teno_) {
        K_ = s0.L_.K;
        M_ = s0.L_.M;
        FQ_ = (int)(s0.L_.S * s0.L_.Q);
        small_ = s0.loc8_ != nullptr;
    }
    for (const auto& c : cases_) {
        if (c->L_.teno != teno_ || (teno_ && (c->L_.K != K_ || c->L_.M != M_ ||
                                              (int)(c->L_.S * c->L_.Q) != FQ_ || (c->loc8_ != nullptr) != small_)))
            fail("ensemble cases must share the reconstruction layout (order, stencil size, tile width)");
    }
    derivable_ = teno_;
    for (const auto& c : cases_) derivable_ = derivable_ && c->cvert_ != nullptr;
    hip_ok(hipStreamCreate(&stream_), "hipStreamCreate");
    hip_ok(hipMalloc(&d_cases_, cases_.size() * sizeof(dev::CaseBufs)), "hipMalloc");
    upload_table();
}

Ensemble::~Ensemble() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (auto& e : prof_) (void)hipEventDestroy(e.second);
    if (d_cases_) (void)hipFree(d_cases_);
    if (d_bar_) (void)hipFree(d_bar_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

uint64_t Ensemble::n_cells() const {
    uint64_t n = 0;
    for (const auto& c : cases_) n += c->L_.N;
    return n;
}

void Ensemble::upload_table() {
    std::vector<dev::CaseBufs> tab;
    tab.reserve(cases_.size());
    const uint32_t bs = tune_.flux_block;
    uint32_t bt = 0, bf = 0;
    gated_ = false;
    split_ = true;
    for (const auto& c : cases_) {
        dev::CaseBufs b = c->case_bufs();
        gated_ = gated_ || b.t.gate != 0;  // the TENO launch then reserves LDS for the network
        split_ = split_ && b.t.gate != 0 && b.t.blk_mixed != nullptr;
        b.blk_teno0 = bt;
        b.blk_flux0 = bf;
        bt += (c->L_.N + 63) / 64;
        bf += (c->L_.N + bs - 1) / bs;
        tab.push_back(b);
    }
    teno_blocks_ = bt;
    flux_blocks_ = bf;
    hip_ok(hipMemcpy(d_cases_, tab.data(), tab.size() * sizeof(dev::CaseBufs), hipMemcpyHostToDevice), "hipMemcpy");
}

void Ensemble::drop_graph() {
    if (exec_) { hip_ok(hipGraphExecDestroy(exec_), "hipGraphExecDestroy"); exec_ = nullptr; }
    if (graph_) { hip_ok(hipGraphDestroy(graph_), "hipGraphDestroy"); graph_ = nullptr; }
}

void Ensemble::set_indicator(const float* params, double threshold) {
    if (persistent_ && params) fail("the persistent step runs plain TENO: switch it off before setting an indicator");
    for (auto& c : cases_) c->set_indicator(params, threshold);
    upload_table();
    drop_graph();
    tuned_ = false;  // the gated kernel has its own best shape
}

void Ensemble::set_indicator_cadence(bool every_stage) {
    gate_per_stage_ = every_stage;
    drop_graph();
}

void Ensemble::set_indicator_split(bool split) {
    for (auto& c : cases_) c->set_indicator_split(split);
    upload_table();
    drop_graph();
    tuned_ = false;
}

void Ensemble::set_tune(const dev::Tune& t) {
    auto bs_ok = [](uint32_t b) { return b == 64 || b == 128 || b == 256; };
    if (!bs_ok(t.flux_block)) fail("block sizes must be 64, 128 or 256 cells");
    if (persistent_ && t.flux_block != 256) fail("the persistent step runs flux and spectral blocks of 256 cells");
    auto occ_ok = [](int o) { return o == 0 || o == 5 || o == 6; };
    if (t.teno_face_basis != 0 && t.teno_face_basis != 1)
        fail("TENO face-basis source must be 0 (stored) or 1 (derived)");
    if (t.teno_face_basis == 1) {
        if (!derivable_)
            fail("derived TENO face basis: p = 2, Legendre basis, default stencil size, tiles of <= 256 cells");
        if (!(t.teno_occ == 0 || t.teno_occ == 4 || t.teno_occ == 5) || !occ_ok(t.teno_occ_reuse))
            fail("derived TENO face basis: occupancy target 0, 4 or 5");
    } else if (!occ_ok(t.teno_occ) || !occ_ok(t.teno_occ_reuse)) {
        fail("TENO occupancy target must be 0, 5 or 6");
    }
    tune_ = t;
    tune_.spec_block = t.flux_block;  // flux and spectral blocks share the case table's offsets
    tuned_ = true;
    upload_table();
    drop_graph();
}

// Times the candidate shapes on the cases' current states with launches that
// write only scratch outputs (face states, the first stage's U and
// accumulator, the radii); gated-cell counters are restored afterwards.
void Ensemble::autotune() {
    if (persistent_) return;  // the persistent step's shapes are fixed
    drop_graph();
    const int n = size();
    hipEvent_t e0, e1;
    hip_ok(hipEventCreate(&e0), "hipEventCreate");
    hip_ok(hipEventCreate(&e1), "hipEventCreate");
    auto timed = [&](const std::function<void()>& f) {
        f();
        hip_ok(hipEventRecord(e0, stream_), "hipEventRecord");
        for (int r = 0; r < 3; ++r) f();
        hip_ok(hipEventRecord(e1, stream_), "hipEventRecord");
        hip_ok(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms;
    };
    dev::Tune best = tune_;
    const int corrected = (teno_ && cfg_.teno.variant == TenoVariant::Corrected) ? 1 : 0;
    const int op = cfg_.integrator == Integrator::RK4 ? dev::kRK4_1 : dev::kSSP1;
    const dev::BatchStage st{op, teno_ ? 0 : 1, corrected, dev::kBufUn, dev::kBufUa, cfg_.positivity ? 1 : 0};
    float bt = 1e30f;
    for (uint32_t bs : {64u, 128u, 256u}) {
        tune_.flux_block = tune_.spec_block = bs;
        upload_table();
        const float t = timed([&] {
            hip_ok(dev::spectral_batch(d_cases_, n, flux_blocks_, ph_, bs, stream_), "spectral_batch");
            hip_ok(dev::flux_batch(d_cases_, n, flux_blocks_, st, (int)cfg_.riemann, ph_, bs, stream_), "flux_batch");
        });
        if (t < bt) { bt = t; best.flux_block = bs; }
    }
    tune_.flux_block = tune_.spec_block = best.flux_block;
    upload_table();
    if (teno_) {
        std::vector<uint32_t*> saved(cases_.size(), nullptr);
        if (gated_) {
            for (size_t k = 0; k < cases_.size(); ++k) {
                const Solver& c = *cases_[k];
                hip_ok(hipMalloc((void**)&saved[k], c.gated_slots_ * sizeof(uint32_t)), "hipMalloc");
                hip_ok(hipMemcpyAsync(saved[k], c.gated_, c.gated_slots_ * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                      stream_), "hipMemcpyAsync");
            }
        }
        const bool occ_tunable = K_ == 6 && M_ == 12 && FQ_ == 3 && small_ && !(gated_ && split_);
        // the plain pass also tries the derived face basis at 0, 4 and 5
        bt = 1e30f;
        for (int fb : {0, 1}) {
            if (fb == 1 && !(derivable_ && !gated_)) continue;
            for (int occ : {0, 4, 5, 6}) {
                if (fb == 0 && (occ == 4 || (occ != 0 && !occ_tunable))) continue;
                if (fb == 1 && occ == 6) continue;
                const float t = timed([&] {
                    hip_ok(dev::teno_batch(d_cases_, n, teno_blocks_, dev::kBufUn, K_, M_, FQ_, small_, gated_, occ,
                                           0, split_, fb == 1, ph_, stream_), "teno_batch");
                });
                if (t < bt) { bt = t; best.teno_occ = occ; best.teno_face_basis = fb; }
            }
        }
        if (gated_ && !gate_per_stage_) {  // passes that reuse the decisions just recorded
            bt = 1e30f;
            for (int occ : {0, 5, 6}) {
                if (occ != 0 && !occ_tunable) continue;
                const float t = timed([&] {
                    hip_ok(dev::teno_batch(d_cases_, n, teno_blocks_, dev::kBufUn, K_, M_, FQ_, small_, gated_, occ,
                                           1, split_, false, ph_, stream_), "teno_batch");
                });
                if (t < bt) { bt = t; best.teno_occ_reuse = occ; }
            }
        }
        if (gated_) {
            for (size_t k = 0; k < cases_.size(); ++k) {
                const Solver& c = *cases_[k];
                hip_ok(hipMemcpyAsync(c.gated_, saved[k], c.gated_slots_ * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                      stream_), "hipMemcpyAsync");
            }
            hip_ok(hipStreamSynchronize(stream_), "hipStreamSynchronize");
            for (uint32_t* p : saved) hip_ok(hipFree(p), "hipFree");
        }
    }
    hip_ok(hipEventDestroy(e0), "hipEventDestroy");
    hip_ok(hipEventDestroy(e1), "hipEventDestroy");
    hip_ok(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    set_tune(best);
}

void Ensemble::mark(int tag) {
    if (!prof_on_) return;
    if (prof_used_ == prof_.size()) {
        hipEvent_t e;
        hip_ok(hipEventCreate(&e), "hipEventCreate");
        prof_.push_back({tag, e});
    }
    prof_[prof_used_].first = tag;
    hip_ok(hipEventRecord(prof_[prof_used_].second, stream_), "hipEventRecord");
    ++prof_used_;
}

void Ensemble::launch_step(bool prof) {
    (void)prof;
    const int n = size();
    const int64_t cap = cases_[0]->dt_hist_cap_;
    const uint32_t bs = tune_.flux_block;
    if (cfg_.use_cfl) hip_ok(dev::spectral_batch(d_cases_, n, flux_blocks_, ph_, bs, stream_), "spectral_batch");
    hip_ok(dev::finalize_batch(d_cases_, n, cfg_.cfl, cfg_.use_cfl ? 0.0 : cfg_.dt, cap, bs, stream_),
           "finalize_batch");
    mark(kStatDtMs);
    const int corrected = (teno_ && cfg_.teno.variant == TenoVariant::Corrected) ? 1 : 0;
    auto stage = [&](int op, int in, int out) {
        if (teno_) {
            const bool first = op == dev::kSSP1 || op == dev::kRK4_1;
            const int reuse = (gated_ && !gate_per_stage_ && !first) ? 1 : 0;
            hip_ok(dev::teno_batch(d_cases_, n, teno_blocks_, in, K_, M_, FQ_, small_, gated_,
                                   reuse ? tune_.teno_occ_reuse : tune_.teno_occ, reuse, split_,
                                   tune_.teno_face_basis == 1 && !gated_, ph_, stream_),
                   "teno_batch");
            mark(kStatTenoMs);
        }
        const dev::BatchStage st{op, teno_ ? 0 : 1, corrected, in, out, cfg_.positivity ? 1 : 0};
        hip_ok(dev::flux_batch(d_cases_, n, flux_blocks_, st, (int)cfg_.riemann, ph_, bs, stream_), "flux_batch");
        mark(kStatFluxMs);
    };
    if (cfg_.integrator == Integrator::SSPRK3) {
        stage(dev::kSSP1, dev::kBufUn, dev::kBufUa);
        stage(dev::kSSP2, dev::kBufUa, dev::kBufUb);
        stage(dev::kSSP3, dev::kBufUb, dev::kBufUn);
    } else {
        stage(dev::kRK4_1, dev::kBufUn, dev::kBufUa);
        stage(dev::kRK4_2, dev::kBufUa, dev::kBufUb);
        stage(dev::kRK4_3, dev::kBufUb, dev::kBufUa);
        stage(dev::kRK4_4, dev::kBufUa, dev::kBufUn);
    }
}

void Ensemble::set_persistent(bool on) {
    if (on == persistent_) return;
    if (on) {
        if (teno_ && !(K_ == 6 && M_ == 12 && FQ_ == 3 && small_))
            fail("the persistent step runs TENO at p = 2 with the default stencil size and tile width");
        if (gated_) fail("the persistent step runs plain TENO: switch the indicator off first");
        if (cfg_.positivity) fail("the persistent step has no positivity limiter");
        int dev = 0, coop = 0;
        hip_ok(hipGetDevice(&dev), "hipGetDevice");
        hip_ok(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev), "hipDeviceGetAttribute");
        if (!coop) fail("this device has no cooperative launches");
        int khz = 0;
        hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
        if (khz <= 0) fail("the device reports no wall-clock rate");
        spin_ticks_ = (long long)(wait_limit_ms_ * khz);
        if (!d_bar_) {
            hip_ok(hipMalloc((void**)&d_bar_, sizeof(dev::StepBarrier)), "hipMalloc");
            hip_ok(hipMemset(d_bar_, 0, sizeof(dev::StepBarrier)), "hipMemset");
        }
        dev::Tune t = tune_;
        t.flux_block = 256;
        persistent_ = true;
        set_tune(t);  // rebuilds the case table's block offsets for 256-cell blocks
    } else {
        persistent_ = false;
        tuned_ = false;  // the per-phase shapes are timed again
        drop_graph();
    }
}

void Ensemble::set_persistent_wait_limit(double ms) {
    if (!(ms >= 0.0) || ms > 10000.0) fail("the wait limit must be within [0, 10000] ms");
    wait_limit_ms_ = ms;
    if (persistent_) {
        int dev = 0, khz = 0;
        hip_ok(hipGetDevice(&dev), "hipGetDevice");
        hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
        spin_ticks_ = (long long)(wait_limit_ms_ * khz);
    }
}

// After a persistent run has synchronised: a block that gave up at a barrier
// leaves the barrier state inconsistent and the states invalid.
void Ensemble::check_barrier() {
    dev::StepBarrier b{};
    hip_ok(hipMemcpy(&b, d_bar_, sizeof(b), hipMemcpyDeviceToHost), "hipMemcpy");
    if (b.err == 0) return;
    hip_ok(hipMemset(d_bar_, 0, sizeof(dev::StepBarrier)), "hipMemset");
    fail("persistent step: a block waited longer than " + std::to_string(wait_limit_ms_) +
         " ms at a grid barrier; the step was abandoned and the cases' states are not valid");
}

// One step of every case: the spectral radii and dt, then per stage the TENO
// and the flux + update phases, in one cooperative launch.
void Ensemble::launch_persistent(long long* clocks) {
    dev::PersistentStep ps{};
    ps.clocks = clocks;
    ps.bar = d_bar_;
    ps.spin_ticks = spin_ticks_;
    if (cfg_.integrator == Integrator::SSPRK3) {
        ps.n_stages = 3;
        const int op[3] = {dev::kSSP1, dev::kSSP2, dev::kSSP3};
        const int in[3] = {dev::kBufUn, dev::kBufUa, dev::kBufUb};
        const int out[3] = {dev::kBufUa, dev::kBufUb, dev::kBufUn};
        for (int k = 0; k < 3; ++k) { ps.op[k] = op[k]; ps.in[k] = in[k]; ps.out[k] = out[k]; }
    } else {
        ps.n_stages = 4;
        const int op[4] = {dev::kRK4_1, dev::kRK4_2, dev::kRK4_3, dev::kRK4_4};
        const int in[4] = {dev::kBufUn, dev::kBufUa, dev::kBufUb, dev::kBufUa};
        const int out[4] = {dev::kBufUa, dev::kBufUb, dev::kBufUa, dev::kBufUn};
        for (int k = 0; k < 4; ++k) { ps.op[k] = op[k]; ps.in[k] = in[k]; ps.out[k] = out[k]; }
    }
    ps.fo = teno_ ? 0 : 1;
    ps.corrected = (teno_ && cfg_.teno.variant == TenoVariant::Corrected) ? 1 : 0;
    ps.use_cfl = cfg_.use_cfl ? 1 : 0;
    ps.cfl = cfg_.cfl;
    ps.fixed_dt = cfg_.use_cfl ? 0.0 : cfg_.dt;
    ps.hist_cap = cases_[0]->dt_hist_cap_;
    hip_ok(dev::ensemble_step(d_cases_, size(), teno_blocks_, flux_blocks_, ps, (int)cfg_.riemann, teno_, ph_, stream_,
                              &persistent_grid_), "ensemble_step");
}

void Ensemble::finish(int64_t s0, int64_t n) {
    for (auto& c : cases_) c->finish_steps(s0, n, nullptr);
}

// Issued in chunks of at most the dt ring's length (every case's ring has
// the same length), each chunk's dt values read back before the next.
void Ensemble::run_steps(int64_t n) {
    if (n <= 0) return;
    check_dt_config(cfg_);
    const int64_t cap = cases_[0]->dt_hist_cap_;
    if (persistent_) {
        for (int64_t done = 0; done < n;) {
            const int64_t m = std::min<int64_t>(n - done, cap);
            const int64_t s0 = cases_[0]->device_step();
            for (int64_t i = 0; i < m; ++i) launch_persistent(nullptr);
            hip_ok(hipStreamSynchronize(stream_), "hipStreamSynchronize");
            check_barrier();
            finish(s0, m);
            done += m;
        }
        return;
    }
    if (!tuned_) autotune();
    if (!exec_) {
        hip_ok(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launch_step(false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_ok(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
        hip_ok(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    for (int64_t done = 0; done < n;) {
        const int64_t m = std::min<int64_t>(n - done, cap);
        const int64_t s0 = cases_[0]->device_step();
        for (int64_t i = 0; i < m; ++i) hip_ok(hipGraphLaunch(exec_, stream_), "hipGraphLaunch");
        hip_ok(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        finish(s0, m);
        done += m;
    }
}

// Persistent mode: every step's phases from the device wall clock stamped
// at the barriers (radii + dt from the start, TENO and flux per stage).
void Ensemble::run_steps_profiled_persistent(int64_t n, double* stats) {
    const int64_t cap = cases_[0]->dt_hist_cap_;
    const int stages = cfg_.integrator == Integrator::SSPRK3 ? 3 : 4, marks = 3 + 2 * stages;
    int dev = 0, khz = 0;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
    if (khz <= 0) fail("the device reports no wall-clock rate");
    const double ms_per_tick = 1.0 / khz;
    const int64_t m_max = std::min<int64_t>(n, cap);
    long long* d_clk = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_ok(hipMalloc((void**)&d_clk, (size_t)m_max * marks * sizeof(long long)), "hipMalloc");
    std::vector<long long> clk((size_t)m_max * marks);
    try {
        hip_ok(hipEventCreate(&e0), "hipEventCreate");
        hip_ok(hipEventCreate(&e1), "hipEventCreate");
        for (int64_t done = 0; done < n;) {
            const int64_t m = std::min<int64_t>(n - done, cap);
            const int64_t s0 = cases_[0]->device_step();
            hip_ok(hipEventRecord(e0, stream_), "hipEventRecord");
            for (int64_t i = 0; i < m; ++i) launch_persistent(d_clk + i * marks);
            hip_ok(hipEventRecord(e1, stream_), "hipEventRecord");
            hip_ok(hipStreamSynchronize(stream_), "hipStreamSynchronize");
            check_barrier();
            float tot = 0.0f;
            hip_ok(hipEventElapsedTime(&tot, e0, e1), "hipEventElapsedTime");
            stats[kStatTotalMs] += tot;
            hip_ok(hipMemcpy(clk.data(), d_clk, (size_t)m * marks * sizeof(long long), hipMemcpyDeviceToHost),
                   "hipMemcpy");
            for (int64_t i = 0; i < m; ++i) {
                const long long* c = clk.data() + i * marks;
                stats[kStatDtMs] += (double)(c[2] - c[0]) * ms_per_tick;
                stats[kStatDtN] += 1.0;
                for (int st = 0; st < stages; ++st) {
                    if (teno_) {
                        stats[kStatTenoMs] += (double)(c[3 + 2 * st] - c[2 + 2 * st]) * ms_per_tick;
                        stats[kStatTenoN] += 1.0;
                    }
                    stats[kStatFluxMs] += (double)(c[4 + 2 * st] - c[3 + 2 * st]) * ms_per_tick;
                    stats[kStatFluxN] += 1.0;
                }
            }
            finish(s0, m);
            done += m;
        }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(d_clk);
        throw;
    }
    hip_ok(hipEventDestroy(e0), "hipEventDestroy");
    hip_ok(hipEventDestroy(e1), "hipEventDestroy");
    hip_ok(hipFree(d_clk), "hipFree");
}

void Ensemble::run_steps_profiled(int64_t n, double* stats) {
    for (int k = 0; k < kStatCount; ++k) stats[k] = 0.0;
    if (n <= 0) return;
    check_dt_config(cfg_);
    if (persistent_) {
        run_steps_profiled_persistent(n, stats);
        return;
    }
    if (!tuned_) autotune();
    const int64_t cap = cases_[0]->dt_hist_cap_;
    for (int64_t done = 0; done < n;) {
        const int64_t m = std::min<int64_t>(n - done, cap);
        const int64_t s0 = cases_[0]->device_step();
        prof_on_ = true;
        prof_used_ = 0;
        try {
            mark(-1);
            for (int64_t i = 0; i < m; ++i) launch_step(true);
        } catch (...) {
            prof_on_ = false;
            throw;
        }
        prof_on_ = false;
        hip_ok(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        for (size_t i = 1; i < prof_used_; ++i) {
            float ms = 0.0f;
            hip_ok(hipEventElapsedTime(&ms, prof_[i - 1].second, prof_[i].second), "hipEventElapsedTime");
            stats[prof_[i].first] += ms;
            stats[prof_[i].first + 1] += 1.0;
        }
        float tot = 0.0f;
        hip_ok(hipEventElapsedTime(&tot, prof_[0].second, prof_[prof_used_ - 1].second), "hipEventElapsedTime");
        stats[kStatTotalMs] += tot;
        finish(s0, m);
        done += m;
    }
}

}  // namespace mlr
