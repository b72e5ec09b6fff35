// lima_engine.hpp — header-only C++ wrapper over lima.h with the reference's
// Engine interface (Quantom/Quantom/Engine.cuh:109-208), for drop-in callers of
// the per-step path:
//
//     lima::Engine engine(systems, params);      // Engine(Simulation*, ForceField), Engine.cu:9-38
//     for (...) { engine.deviceMaster();         // Engine.cu:44-47  -> queues one step
//                 engine.hostMaster(); }         // Engine.cu:49-75  -> waits once per list period
//     engine.sync();                             // waits for queued steps, surfaces errors
//     engine.timings;                            // Engine.cuh:123 (us: device, rebuild, host)
//
// deviceMaster() queues one step without waiting: a replay of the captured graph of that
// step (the whole rebuild cycle when n steps are queued at once by step(n)).  The
// thermostat and the scheduled list rebuild that the reference's hostMaster() runs on the
// host run on the device inside the step, so hostMaster() only has to wait for the device
// once every sync_every() calls (the list interval) and report device errors.  An error
// raised by a step therefore surfaces at most one list period later; every call that
// reads results (energies, state, parameters) waits for the queued steps first.
//
// Errors throw lima::Error carrying the lima.h status code instead of exit().
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "lima.h"

namespace lima {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Int3 {
    long long x = 0, y = 0, z = 0;
};

inline lima_run_params default_params() {
    lima_run_params p{};
    p.cutoff = 0.6f;
    p.nlist_interval = 10;
    p.thermostat = LIMA_THERMOSTAT_BERENDSEN;
    p.thermostat_tau = 1e-4;
    p.accumulation = LIMA_ACCUM_FP32;
    p.enable_solvents = 1;
    p.enable_dihedrals = 1;
    p.max_compound_neighbors = 256;
    p.max_compound_solvents = 16384;
    p.max_solvent_neighbors = 512;
    p.device = 0;
    return p;
}

class Engine {
public:
    Engine(const std::vector<lima_system>& systems, const lima_run_params& params)
        : n_systems_((int)systems.size()),
          sync_every_(params.nlist_interval > 0 ? params.nlist_interval : 10) {
        check(lima_create(systems.data(), (int32_t)systems.size(), &params, &h_), nullptr);
    }
    Engine(const lima_system& system, const lima_run_params& params) : Engine(std::vector<lima_system>{system}, params) {}
    ~Engine() { lima_destroy(h_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void deviceMaster() {
        check(lima_step_async(h_, 1));
        queued_++;
    }
    void hostMaster() {
        if (queued_ >= sync_every_) sync();
    }
    // waits for every queued step, reports device errors and refreshes timings
    void sync() {
        queued_ = 0;
        check(lima_sync(h_));
        double t[5];
        check(lima_get_timings(h_, t));
        timings.x = (long long)t[0];
        timings.y = (long long)t[1];
        timings.z = (long long)t[2];
    }
    int sync_every() const { return sync_every_; }
    int systems() const { return n_systems_; }
    void step(int n) {
        queued_ = 0;
        check(lima_step(h_, n));
    }
    void rebuildNeighborLists() { check(lima_rebuild_neighbors(h_)); }
    std::vector<lima_energies> energies() {
        std::vector<lima_energies> e(n_systems_);
        check(lima_get_energies(h_, e.data()));
        return e;
    }
    // energy series of Analyzer::analyzeEnergy, summed in the step (records on step 0, every, 2*every, ...)
    std::vector<lima_energies> stepObserved(int n_steps, int every) {
        if (n_steps < 0 || every <= 0) throw Error(LIMA_E_ARG, "n_steps < 0 or every <= 0");
        std::vector<lima_energies> e((size_t)((n_steps + every - 1) / every) * n_systems_);
        check(lima_step_observed(h_, n_steps, every, e.data()));
        return e;
    }
    // parameter vector of one system: sigma[32], epsilon[32], kb, b0, k_theta, theta_0
    std::vector<double> parameters(int system) {
        std::vector<double> v(LIMA_N_PARAMETERS);
        check(lima_get_parameters(h_, system, v.data()));
        return v;
    }
    void setParameters(int system, const std::vector<double>& v) {
        if (v.size() != LIMA_N_PARAMETERS) throw Error(LIMA_E_ARG, "expected LIMA_N_PARAMETERS values");
        check(lima_set_parameters(h_, system, v.data()));
    }
    // loss per system and its gradient [systems()][LIMA_N_PARAMETERS]; the state is left as it was
    std::vector<double> adjoint(const lima_adjoint_spec& spec, std::vector<double>* grad) {
        std::vector<double> loss(n_systems_);
        if (grad) grad->assign((size_t)n_systems_ * LIMA_N_PARAMETERS, 0.0);
        check(lima_adjoint_run(h_, &spec, loss.data(), grad ? grad->data() : nullptr));
        return loss;
    }
    lima_engine* handle() { return h_; }

    Int3 timings;

private:
    lima_engine* h_ = nullptr;
    int n_systems_ = 0;
    int sync_every_ = 10;
    int queued_ = 0;
    void check(int st) { check(st, h_); }
    static void check(int st, lima_engine* h) {
        if (st != LIMA_OK) throw Error(st, lima_last_error(h));
    }
};

}  // namespace lima
