// Job server for C++ callers: a chain and a program (Relax, CoolDown,
// HysteresisLoop) submitted to the engine's continuous-batching server
// (vg_server_* in vegas_hip.h), with each measured stage's StatSensor values
// (T, |h|, <E>, C, <|M|>, chi, U4; instrument.rs:113-123) as the result.
#pragma once

#include <vector>

#include "machine.hpp"
#include "program.hpp"

namespace vegas {

// The stages a program runs on a machine whose field starts at field0:
// stages that do not set the field keep the current one (Relax and CoolDown
// only change the temperature, program.rs:97-212).
template <class P>
std::vector<vg_stage> stages_of(const P& program, double field0 = 0.0) {
    std::vector<vg_stage> out;
    double field = field0;
    for (const Stage& s : program.schedule()) {
        if (s.sets_field) field = s.field;
        vg_stage v{};
        v.temperature = s.temperature;
        v.field = field;
        v.field_dir[2] = 1.0;  // Field::new(S::up(), field)
        v.relax = s.relax;
        v.measure = s.relax_only ? 0 : s.measure;
        out.push_back(v);
    }
    return out;
}

struct JobServerConfig {
    uint64_t memory_bytes = 1ull << 30;  // device pool for the running jobs' spins
    uint32_t max_jobs = 256;             // jobs on the device at once
    uint32_t quantum_sweeps = 100;       // sweeps per scheduling quantum
    uint32_t tile = 2;                   // Ising sc tile
    bool graph = false;                  // replay quanta from hipGraphs
    uint64_t sweep_budget = 0;           // per-sweep work cap, shortest jobs first (0: none)
};

class JobServer {
  public:
    explicit JobServer(const JobServerConfig& c = JobServerConfig(),
                       vg_ctx* ctx = default_context()) {
        vg_server_config cfg{};
        cfg.memory_bytes = c.memory_bytes;
        cfg.max_jobs = c.max_jobs;
        cfg.quantum_sweeps = c.quantum_sweeps;
        cfg.tile = c.tile;
        cfg.use_graph = c.graph ? 1 : 0;
        cfg.sweep_budget = c.sweep_budget;
        check(vg_server_create(ctx, &cfg, &server_));
    }
    ~JobServer() { vg_server_destroy(server_); }
    JobServer(const JobServer&) = delete;
    JobServer& operator=(const JobServer&) = delete;

    // Submits the chain of Hamiltonian h (its lattice from the Exchange term)
    // with integrator kind I running `program`; the spins start from
    // State::rand_with_size drawn from `seed`.
    template <class S, class I, class H, class P>
    uint64_t submit(const H& h, const I&, const P& program, uint64_t seed, size_t n_sites = 0) {
        if (I::proposal == VG_PROPOSAL_WOLFF)
            throw Error(ErrorKind::NotImplemented,
                        "the job server runs Metropolis chains; Wolff chains run in a Machine");
        vg_job_desc job{};
        job.chain.lattice = lattice_of(h, n_sites);
        job.chain.spin = S::kind;
        job.chain.proposal = I::proposal;
        job.chain.hamiltonian = describe(h);
        job.chain.seed = seed;
        const std::vector<vg_stage> stages = stages_of(program);
        job.n_stages = (uint32_t)stages.size();
        job.stages = stages.data();
        uint64_t id = 0;
        check(vg_server_submit(server_, &job, &id));
        n_stages_[id] = job.n_stages;
        return id;
    }
    vg_job_status wait(uint64_t id, double timeout_s = -1.0) {
        vg_job_status st{};
        check(vg_server_wait(server_, id, timeout_s, &st));
        return st;
    }
    vg_job_status status(uint64_t id) {
        vg_job_status st{};
        check(vg_server_status(server_, id, &st));
        return st;
    }
    std::vector<vg_stage_result> results(uint64_t id) {
        std::vector<vg_stage_result> r(n_stages_.at(id));
        check(vg_server_results(server_, id, r.data(), (uint32_t)r.size()));
        return r;
    }
    void cancel(uint64_t id) { check(vg_server_cancel(server_, id)); }
    vg_server_stats stats() {
        vg_server_stats s{};
        check(vg_server_get_stats(server_, &s));
        return s;
    }

  private:
    vg_server* server_ = nullptr;
    std::map<uint64_t, uint32_t> n_stages_;
};

}  // namespace vegas
