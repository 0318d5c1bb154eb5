// Job server: continuous batching of simulation jobs on one GPU.
//
// A worker thread owns a private HIP stream and a working batch whose
// membership changes as jobs come and go.  Every quantum it
//   1. retires cancelled jobs and admits queued ones, in arrival order, while
//      a slot is free and the job's spins fit in the device pool (first-fit
//      over one allocation made at start: admission by HBM footprint);
//   2. initialises newly admitted chains on the device (Philox random start);
//   3. runs one sweep block over every running job -- one launch per colour
//      pass and kernel group covers all of them -- where each job runs
//      min(quantum, sweeps left in its current stage) sweeps (the pass
//      kernels skip a chain past its limit), so every stage boundary falls
//      exactly on its sweep count;
//   4. turns each job's per-sweep integer sums into E and |M| with the
//      reference's conventions and feeds the stage accumulators in sweep
//      order; finished jobs release their pool space.
// A job's trajectory depends only on its seed, its sweep counter and its own
// thermostat, and its sums are exact integers, so its results are the same
// bit for bit alone or in any mix of jobs.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <iterator>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "engine.hpp"

using vg::ChainDev;

namespace vgi {
namespace {

// Relative cost of a Heisenberg site update to an Ising one (pass times per
// site on the MI355X: about 11 ns vs 0.75 ns per 1000 sites).
constexpr uint64_t kHeisSiteWeight = 15;

// First-fit allocator of byte ranges in one device pool; blocks are
// 256-byte aligned and neighbours coalesce on release.
class PoolAllocator {
  public:
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(uint64_t size, uint64_t& off) {
        size = round_up(size, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < size) continue;
            off = it->first;
            const uint64_t rest = it->second - size;
            free_.erase(it);
            if (rest) free_[off + size] = rest;
            used_ += size;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t size) {
        size = round_up(size, 256);
        used_ -= size;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += size;
                return;
            }
        }
        free_[off] = size;
    }
    uint64_t used() const { return used_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0;
};

// Accumulator of accumulator.rs:23-63 for E and |M| (E needs no 4th power).
struct StageAcc {
    double e = 0, e2 = 0, m = 0, m2 = 0, m4 = 0;
    uint64_t n = 0;
    void collect(double ev, double mv) {
        e += ev;
        e2 += ev * ev;
        m += mv;
        m2 += mv * mv;
        m4 += mv * mv * mv * mv;
        ++n;
    }
};

struct Job {
    uint64_t id = 0;
    ChainHost ch;
    ChainDev dev;
    std::vector<vg_stage> stages;
    std::vector<vg_stage_result> results;
    int state = VG_JOB_QUEUED;
    bool cancel = false;
    bool fresh = true;  // admitted, spins not initialised yet
    uint64_t pool_off = 0, bytes = 0;
    uint32_t stage = 0;
    uint64_t in_stage = 0;  // sweeps done in the current stage
    uint64_t sweeps_total = 0, sweeps_done = 0;
    StageAcc acc;
    double last_e = 0.0, last_m = 0.0;
    double submit_s = -1.0, start_s = -1.0, end_s = -1.0;
    std::string error;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Thermostat of a stage (thermostat.rs:29-34 clamp, |field| on read).
void enter_stage(Job& j) {
    const vg_stage& st = j.stages[j.stage];
    j.ch.temperature = st.temperature < DBL_EPSILON ? DBL_EPSILON : st.temperature;
    j.ch.field_mag = st.field;
    for (int k = 0; k < 3; ++k) j.ch.field_dir[k] = st.field_dir[k];
    if (is_ising(j.ch.group))
        ising_thresholds(j.ch, j.dev);
    else
        heis_params(j.ch, j.dev);
    j.in_stage = 0;
    j.acc = StageAcc();
}

// StatSensor's line (instrument.rs:113-123) from the stage accumulators.
vg_stage_result finish_stage(const Job& j) {
    vg_stage_result r{};
    const double T = j.ch.temperature, n = (double)j.ch.n_sites;
    r.temperature = T;
    r.field = std::fabs(j.ch.field_mag);
    r.measured = j.acc.n;
    if (j.acc.n) {
        const double c = (double)j.acc.n;
        const double me = j.acc.e / c, mm = j.acc.m / c;
        r.energy = me;
        r.specific_heat = (j.acc.e2 / c - me * me) / (n * (T * T));
        r.magnetization = mm;
        r.susceptibility = (j.acc.m2 / c - mm * mm) / (n * T);
        const double m2 = j.acc.m2 / c;
        r.binder = 1.0 - (j.acc.m4 / c) / (3.0 * (m2 * m2));
    }
    return r;
}

}  // namespace
}  // namespace vgi

struct vg_server {
    vg_server_config cfg{};
    vg_ctx own_ctx;  // device + the worker's private stream
    void* pool = nullptr;
    vgi::PoolAllocator alloc;
    uint64_t pool_peak = 0;
    vg_batch B;  // working batch over the running jobs
    int64_t* d_records = nullptr;
    int64_t* h_records = nullptr;
    uint64_t records_cap = 0;  // int64 words
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, std::unique_ptr<vgi::Job>> jobs;
    std::vector<vgi::Job*> running;
    bool membership_changed = true;
    bool stop = false;
    bool broken = false;
    std::string broken_msg;
    uint64_t next_id = 1;
    vg_server_stats stats{};
    std::chrono::steady_clock::time_point t0;
    std::thread worker;
};

namespace vgi {
namespace {

uint64_t job_bytes(const Job& j) {
    const int g = j.ch.group;
    return j.dev.ncol * round_up(colour_elems(g, j.dev) * elem_bytes(g), 256);
}

// Makes the working batch's groups the running jobs and sizes its buffers.
int rebuild(vg_server* S) {
    vg_batch& B = S->B;
    B.chains.clear();
    for (Group& G : B.groups) {
        G.members.clear();
        G.host.clear();
    }
    for (size_t i = 0; i < S->running.size(); ++i) {
        Job& j = *S->running[i];
        j.ch.index = (uint32_t)B.groups[j.ch.group].members.size();
        j.dev.slot = (uint32_t)i;
        B.groups[j.ch.group].members.push_back((uint32_t)i);
        B.groups[j.ch.group].host.push_back(j.dev);
        B.chains.push_back(j.ch);
    }
    const uint32_t Q = S->cfg.quantum_sweeps;
    for (int g = 0; g < kNumGroups; ++g) {
        Group& G = B.groups[g];
        if (G.members.empty()) {
            G.total_blocks = 0;
            continue;
        }
        if (G.host.size() > G.chains_cap) {
            if (G.d_chains) quiet(hipFree(G.d_chains));
            if (G.h_staging) quiet(hipHostFree(G.h_staging));
            G.d_chains = nullptr;
            G.h_staging = nullptr;
            G.chains_cap = 0;
            const uint32_t cap = (uint32_t)std::max<size_t>(16, G.host.size() * 2);
            VG_HIP(hipMalloc(&G.d_chains, cap * sizeof(ChainDev)));
            VG_HIP(hipHostMalloc(&G.h_staging, cap * sizeof(ChainDev)));
            G.chains_cap = cap;
        }
        int rc = layout_work(&B, g, is_zmarch(g) ? S->cfg.tile : 1);
        if (rc) return rc;
        // layout_work recorded each chain's blocks in G.host; keep them
        for (size_t k = 0; k < G.members.size(); ++k) S->running[G.members[k]]->dev = G.host[k];
        const uint64_t need = G.partial_bytes_per_sweep * Q;
        if (need > G.partial_cap) {
            if (G.partials) quiet(hipFree(G.partials));
            G.partials = nullptr;
            G.partial_cap = 0;
            const uint64_t cap = std::max<uint64_t>(need + need / 2, 1 << 20);
            VG_HIP(hipMalloc(&G.partials, cap));
            G.partial_cap = cap;
        }
    }
    const uint64_t words = (uint64_t)S->running.size() * Q * 5;
    if (words > S->records_cap) {
        if (S->d_records) quiet(hipFree(S->d_records));
        if (S->h_records) quiet(hipHostFree(S->h_records));
        S->d_records = nullptr;
        S->h_records = nullptr;
        S->records_cap = 0;
        const uint64_t cap = std::max<uint64_t>(words * 2, 4096);
        VG_HIP(hipMalloc(&S->d_records, cap * sizeof(int64_t)));
        VG_HIP(hipHostMalloc(&S->h_records, cap * sizeof(int64_t)));
        S->records_cap = cap;
    }
    B.d_records = S->d_records;
    B.h_records = S->h_records;
    return VG_OK;
}

// One quantum over the running jobs, launched and waited for without the
// lock (only the worker changes the running set and the jobs' device
// state).  Returns the block length in nsw.
int quantum(vg_server* S, uint32_t& nsw_out) {
    vg_batch& B = S->B;
    hipStream_t st = B.ctx->stream;
    const uint32_t Q = S->cfg.quantum_sweeps;
    // With a budget, jobs with the least remaining work run first: they are
    // taken in that order until their per-sweep cost fills the budget (the
    // first always runs), the others skip this quantum (limit 0).
    std::vector<Job*> order(S->running);
    const uint64_t budget = S->cfg.sweep_budget;
    auto cost = [](const Job* j) {
        return j->ch.n_sites * (is_ising(j->ch.group) ? 1u : kHeisSiteWeight);
    };
    if (budget)
        std::stable_sort(order.begin(), order.end(), [&](const Job* a, const Job* b) {
            return (a->sweeps_total - a->sweeps_done) * cost(a) <
                   (b->sweeps_total - b->sweeps_done) * cost(b);
        });
    uint32_t nsw = 0;
    uint64_t used = 0;
    for (Job* j : order) {
        const vg_stage& sg = j->stages[j->stage];
        const uint64_t left = sg.relax + sg.measure - j->in_stage;
        const bool take = budget == 0 || used == 0 || used + cost(j) <= budget;
        if (take) used += cost(j);
        j->dev.limit = take ? (uint32_t)std::min<uint64_t>(left, Q) : 0u;
        j->dev.sweep0 = j->ch.sweeps;
        nsw = std::max(nsw, j->dev.limit);
    }
    for (int g = 0; g < kNumGroups; ++g) {
        Group& G = B.groups[g];
        for (size_t k = 0; k < G.members.size(); ++k) G.host[k] = S->running[G.members[k]]->dev;
        int rc = upload_descriptors(&B, g);
        if (rc) return rc;
    }
    // Chains admitted since the last quantum start from State::rand_with_size.
    for (Job* j : S->running) {
        if (!j->fresh) continue;
        Group& G = B.groups[j->ch.group];
        const ChainDev* dp = G.d_chains + j->ch.index;
        if (is_ising(j->ch.group)) {
            launch_ising_init(j->ch.group, G, j->dev, dp, st);
        } else {
            const uint64_t nc = 3ull * j->dev.ncol * j->dev.plane;
            const uint64_t ni = (uint64_t)j->dev.ncol * j->dev.rows * j->dev.w;
            hipLaunchKernelGGL(vg::heis_clear, dim3((uint32_t)((nc + 255) / 256)), dim3(256), 0, st,
                               dp, (float*)G.arena);
            hipLaunchKernelGGL(vg::heis_init, dim3((uint32_t)((ni + 255) / 256)), dim3(256), 0, st,
                               dp, (float*)G.arena);
        }
        j->fresh = false;
    }
    VG_HIP(hipGetLastError());
    B.use_graph = S->cfg.use_graph != 0;
    VG_HIP(hipEventRecord(S->ev0, st));
    int rc = launch_block(&B, nsw);
    if (rc) return rc;
    VG_HIP(hipEventRecord(S->ev1, st));
    const uint64_t nrec = (uint64_t)S->running.size() * nsw * 5;
    VG_HIP(hipMemcpyAsync(S->h_records, S->d_records, nrec * sizeof(int64_t),
                          hipMemcpyDeviceToHost, st));
    VG_HIP(hipStreamSynchronize(st));
    float ms = 0.0f;
    VG_HIP(hipEventElapsedTime(&ms, S->ev0, S->ev1));
    S->stats.device_ms += ms;
    nsw_out = nsw;
    return VG_OK;
}

// Per-sweep sums -> E and |M| -> stage accumulators, in sweep order (under
// the lock: status readers see consistent progress).
void absorb(vg_server* S, uint32_t nsw, uint64_t& attempts) {
    // The records of chain i, sweep s are at (i * nsw + s) * 5.
    attempts = 0;
    for (size_t i = 0; i < S->running.size(); ++i) {
        Job& j = *S->running[i];
        for (uint32_t s = 0; s < j.dev.limit; ++s) {
            const int64_t* r = S->h_records + (i * nsw + s) * 5u;
            vg_sweep_sums v;
            v.bond = r[0];
            v.mag[0] = r[1];
            v.mag[1] = r[2];
            v.mag[2] = r[3];
            v.axis2 = j.ch.desc.spin == VG_SPIN_HEISENBERG ? r[4] : (int64_t)j.ch.n_sites;
            double e = 0.0, m = 0.0;
            observables(j.ch, v, &e, &m);
            if (j.in_stage >= j.stages[j.stage].relax) j.acc.collect(e, m);
            j.last_e = e;
            j.last_m = m;
            ++j.in_stage;
        }
        j.ch.sweeps += j.dev.limit;
        j.sweeps_done += j.dev.limit;
        attempts += j.ch.n_sites * j.dev.limit;
    }
}

// Stage bookkeeping after a quantum (under the lock): completed stages give
// their result and the next stage's thermostat; finished jobs leave.
void advance(vg_server* S) {
    const double t = now_s(S->t0);
    std::vector<Job*> keep;
    for (Job* j : S->running) {
        bool done = false;
        while (true) {
            const vg_stage& sg = j->stages[j->stage];
            if (j->in_stage < sg.relax + sg.measure) break;
            j->results.push_back(finish_stage(*j));
            if (j->stage + 1 == j->stages.size()) {
                done = true;
                break;
            }
            ++j->stage;
            enter_stage(*j);
        }
        if (done) {
            j->state = VG_JOB_DONE;
            j->end_s = t;
            S->alloc.release(j->pool_off, j->bytes);
            S->stats.completed += 1;
            S->membership_changed = true;
        } else {
            keep.push_back(j);
        }
    }
    S->running.swap(keep);
}

void fail_all(vg_server* S, const std::string& msg) {
    const double t = now_s(S->t0);
    S->broken = true;
    S->broken_msg = msg;
    for (Job* j : S->running) {
        j->state = VG_JOB_FAILED;
        j->error = msg;
        j->end_s = t;
    }
    S->running.clear();
    for (uint64_t id : S->queue) {
        Job& j = *S->jobs[id];
        j.state = VG_JOB_FAILED;
        j.error = msg;
        j.end_s = t;
    }
    S->queue.clear();
}

void worker_main(vg_server* S) {
    if (hipSetDevice(S->own_ctx.device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(S->mu);
        fail_all(S, "hipSetDevice failed in the server worker");
        S->cv_done.notify_all();
        return;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        // Cancellations of running jobs.
        {
            std::vector<Job*> keep;
            const double t = now_s(S->t0);
            for (Job* j : S->running) {
                if (j->cancel || S->stop) {
                    j->state = VG_JOB_CANCELLED;
                    j->end_s = t;
                    S->alloc.release(j->pool_off, j->bytes);
                    S->membership_changed = true;
                } else {
                    keep.push_back(j);
                }
            }
            S->running.swap(keep);
        }
        if (S->stop) break;
        // Admission: arrival order, a free slot and room in the pool.
        while (!S->queue.empty() && S->running.size() < S->cfg.max_jobs) {
            Job& j = *S->jobs[S->queue.front()];
            if (j.cancel) {
                j.state = VG_JOB_CANCELLED;
                j.end_s = now_s(S->t0);
                S->queue.pop_front();
                continue;
            }
            uint64_t off = 0;
            if (!S->alloc.alloc(j.bytes, off)) break;  // head of line waits for memory
            S->queue.pop_front();
            j.pool_off = off;
            const int g = j.ch.group;
            const uint64_t e = elem_bytes(g);
            for (uint32_t c = 0; c < j.dev.ncol; ++c)
                j.dev.off[c] = (off + c * round_up(colour_elems(g, j.dev) * e, 256)) / e;
            j.state = VG_JOB_RUNNING;
            j.start_s = now_s(S->t0);
            j.fresh = true;
            enter_stage(j);
            S->running.push_back(&j);
            S->membership_changed = true;
            S->pool_peak = std::max(S->pool_peak, S->alloc.used());
            S->stats.peak_jobs = std::max<uint32_t>(S->stats.peak_jobs, (uint32_t)S->running.size());
        }
        S->cv_done.notify_all();
        if (S->running.empty()) {
            S->cv_work.wait(lk);
            continue;
        }
        const bool changed = S->membership_changed;
        S->membership_changed = false;
        lk.unlock();
        int rc = changed ? rebuild(S) : VG_OK;
        uint32_t nsw = 0;
        if (rc == VG_OK) rc = quantum(S, nsw);
        const std::string msg = rc == VG_OK ? std::string() : g_error;
        lk.lock();
        if (rc != VG_OK) {
            fail_all(S, msg);
            break;
        }
        uint64_t attempts = 0;
        absorb(S, nsw, attempts);
        S->stats.attempts += attempts;
        S->stats.quanta += 1;
        advance(S);
        S->cv_done.notify_all();
    }
    S->cv_done.notify_all();
}

int validate_job(const vg_job_desc& j) {
    int rc = validate_desc(j.chain);
    if (rc) return rc;
    if (j.chain.lattice.kind == VG_LATTICE_GRAPH)
        return fail(VG_ERR_NOT_IMPLEMENTED,
                    "the job server runs lattice chains; graph chains run in batches");
    if (j.chain.proposal == VG_PROPOSAL_WOLFF)
        return fail(VG_ERR_NOT_IMPLEMENTED, "the job server runs Metropolis chains; Wolff chains run in batches");
    if (j.n_stages == 0 || !j.stages) return fail(VG_ERR_NO_STEPS, "a job needs at least one stage");
    // The job's sweep counter keys Philox with 56 bits, as in a batch run.
    const uint64_t kMaxSweeps = 1ull << 56;
    uint64_t total = 0;
    for (uint32_t k = 0; k < j.n_stages; ++k) {
        const vg_stage& s = j.stages[k];
        if (s.relax + s.measure == 0) return fail(VG_ERR_NO_STEPS, "stage %u has no sweeps", k);
        if (std::isnan(s.temperature) || std::isnan(s.field))
            return fail(VG_ERR_INVALID_ARGUMENT, "stage %u: NaN thermostat parameter", k);
        if (s.relax > kMaxSweeps || s.measure > kMaxSweeps - s.relax ||
            s.relax + s.measure > kMaxSweeps - total)
            return fail(VG_ERR_INVALID_ARGUMENT,
                        "stage %u: the job's sweeps would pass 2^56 (the sweep counter's range)", k);
        total += s.relax + s.measure;
    }
    return VG_OK;
}

}  // namespace
}  // namespace vgi

using namespace vgi;

extern "C" {

int vg_server_create(vg_ctx* ctx, const vg_server_config* config, vg_server** out) {
    if (!ctx || !config || !out) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const vg_server_config& c = *config;
    if (c.memory_bytes < 4096) return fail(VG_ERR_INVALID_ARGUMENT, "memory_bytes too small");
    if (c.max_jobs == 0) return fail(VG_ERR_INVALID_ARGUMENT, "max_jobs must be positive");
    if (c.quantum_sweeps == 0 || c.quantum_sweeps > 256)
        return fail(VG_ERR_INVALID_ARGUMENT, "quantum_sweeps must be in 1..256");
    if (c.tile != 1 && c.tile != 2 && c.tile != 4 && c.tile != 8)
        return fail(VG_ERR_INVALID_ARGUMENT, "tile must be 1, 2, 4 or 8");
    VG_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<vg_server> S(new vg_server);
    S->cfg = c;
    S->own_ctx.device = ctx->device;
    VG_HIP(hipStreamCreateWithFlags(&S->own_ctx.stream, hipStreamNonBlocking));
    S->B.ctx = &S->own_ctx;
    hipError_t e = hipMalloc(&S->pool, c.memory_bytes);
    if (e != hipSuccess) {
        quiet(hipStreamDestroy(S->own_ctx.stream));
        return fail(VG_ERR_OUT_OF_MEMORY, "allocating the job pool (%llu bytes): %s",
                    (unsigned long long)c.memory_bytes, hipGetErrorString(e));
    }
    S->alloc.reset(c.memory_bytes & ~255ull);
    for (Group& G : S->B.groups) {
        G.arena = S->pool;
        G.shared_arena = true;
    }
    if ((e = hipEventCreate(&S->ev0)) != hipSuccess || (e = hipEventCreate(&S->ev1)) != hipSuccess) {
        quiet(hipFree(S->pool));
        quiet(hipStreamDestroy(S->own_ctx.stream));
        return fail(VG_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e));
    }
    S->stats.pool_bytes = c.memory_bytes;
    S->t0 = std::chrono::steady_clock::now();
    vg_server* raw = S.release();
    raw->worker = std::thread(worker_main, raw);
    *out = raw;
    return VG_OK;
}

int vg_server_destroy(vg_server* S) {
    if (!S) return VG_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        const double t = now_s(S->t0);
        for (uint64_t id : S->queue) {
            Job& j = *S->jobs[id];
            j.state = VG_JOB_CANCELLED;
            j.end_s = t;
        }
        S->queue.clear();
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    quiet(hipSetDevice(S->own_ctx.device));
    quiet(hipStreamSynchronize(S->own_ctx.stream));
    for (auto& kv : S->B.graphs) quiet(hipGraphExecDestroy(kv.second));
    S->B.graphs.clear();
    for (Group& G : S->B.groups) {
        if (G.d_chains) quiet(hipFree(G.d_chains));
        if (G.h_staging) quiet(hipHostFree(G.h_staging));
        if (G.d_block_chain) quiet(hipFree(G.d_block_chain));
        if (G.partials) quiet(hipFree(G.partials));
        G.d_chains = nullptr;
        G.h_staging = nullptr;
        G.d_block_chain = nullptr;
        G.partials = nullptr;
        G.arena = nullptr;
    }
    S->B.d_records = nullptr;
    S->B.h_records = nullptr;
    if (S->d_records) quiet(hipFree(S->d_records));
    if (S->h_records) quiet(hipHostFree(S->h_records));
    if (S->pool) quiet(hipFree(S->pool));
    if (S->ev0) quiet(hipEventDestroy(S->ev0));
    if (S->ev1) quiet(hipEventDestroy(S->ev1));
    quiet(hipStreamDestroy(S->own_ctx.stream));
    delete S;
    return VG_OK;
}

int vg_server_submit(vg_server* S, const vg_job_desc* job, uint64_t* job_id) {
    if (!S || !job || !job_id) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_job(*job);
    if (rc) return rc;
    std::unique_ptr<Job> j(new Job);
    j->ch.desc = job->chain;
    const vg_lattice& L = job->chain.lattice;
    j->ch.n_sites = sites_of(L);
    j->ch.group = group_of(job->chain);
    rc = init_chain_dev(j->ch, j->ch.group, 0, j->dev);
    if (rc) return rc;
    j->bytes = job_bytes(*j);
    if (j->bytes > (S->cfg.memory_bytes & ~255ull))
        return fail(VG_ERR_ADMISSION, "job needs %llu bytes of spins; the pool holds %llu",
                    (unsigned long long)j->bytes, (unsigned long long)S->cfg.memory_bytes);
    j->stages.assign(job->stages, job->stages + job->n_stages);
    for (const vg_stage& s : j->stages) j->sweeps_total += s.relax + s.measure;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->broken) return fail(VG_ERR_DEVICE, "server stopped: %s", S->broken_msg.c_str());
    if (S->stop) return fail(VG_ERR_CANCELLED, "server is shutting down");
    j->id = S->next_id++;
    j->submit_s = now_s(S->t0);
    *job_id = j->id;
    S->queue.push_back(j->id);
    S->jobs[j->id] = std::move(j);
    S->stats.submitted += 1;
    S->cv_work.notify_all();
    return VG_OK;
}

namespace {
void fill_status(const Job& j, vg_job_status* st) {
    st->state = j.state;
    st->stage = j.stage;
    st->sweeps_done = j.sweeps_done;
    st->sweeps_total = j.sweeps_total;
    st->submit_s = j.submit_s;
    st->start_s = j.start_s;
    st->end_s = j.end_s;
    st->last_energy = j.last_e;
    st->last_magnetization = j.last_m;
}
bool finished(int state) {
    return state == VG_JOB_DONE || state == VG_JOB_FAILED || state == VG_JOB_CANCELLED;
}
}  // namespace

int vg_server_status(vg_server* S, uint64_t id, vg_job_status* st) {
    if (!S || !st) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return fail(VG_ERR_NOT_FOUND, "no job %llu", (unsigned long long)id);
    fill_status(*it->second, st);
    return VG_OK;
}

int vg_server_wait(vg_server* S, uint64_t id, double timeout_s, vg_job_status* st) {
    if (!S) return fail(VG_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return fail(VG_ERR_NOT_FOUND, "no job %llu", (unsigned long long)id);
    Job* j = it->second.get();
    auto pred = [&] { return finished(j->state); };
    if (timeout_s < 0)
        S->cv_done.wait(lk, pred);
    else
        S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    if (st) fill_status(*j, st);
    return VG_OK;
}

int vg_server_results(vg_server* S, uint64_t id, vg_stage_result* out, uint32_t n) {
    if (!S || !out) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return fail(VG_ERR_NOT_FOUND, "no job %llu", (unsigned long long)id);
    const Job& j = *it->second;
    if (j.state == VG_JOB_FAILED) return fail(VG_ERR_DEVICE, "job failed: %s", j.error.c_str());
    if (j.state == VG_JOB_CANCELLED) return fail(VG_ERR_CANCELLED, "job was cancelled");
    if (j.state != VG_JOB_DONE) return fail(VG_ERR_INVALID_ARGUMENT, "job has not finished");
    if (n != j.results.size())
        return fail(VG_ERR_INVALID_ARGUMENT, "job has %zu stages, asked for %u", j.results.size(), n);
    std::copy(j.results.begin(), j.results.end(), out);
    return VG_OK;
}

int vg_server_cancel(vg_server* S, uint64_t id) {
    if (!S) return fail(VG_ERR_INVALID_ARGUMENT, "null server");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->jobs.find(id);
        if (it == S->jobs.end()) return fail(VG_ERR_NOT_FOUND, "no job %llu", (unsigned long long)id);
        it->second->cancel = true;
    }
    S->cv_work.notify_all();
    return VG_OK;
}

int vg_server_release(vg_server* S, uint64_t id) {
    if (!S) return fail(VG_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return fail(VG_ERR_NOT_FOUND, "no job %llu", (unsigned long long)id);
    if (!finished(it->second->state))
        return fail(VG_ERR_INVALID_ARGUMENT, "job %llu has not finished", (unsigned long long)id);
    S->jobs.erase(it);
    return VG_OK;
}

int vg_server_get_stats(vg_server* S, vg_server_stats* st) {
    if (!S || !st) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *st = S->stats;
    st->running = S->running.size();
    st->queued = S->queue.size();
    st->pool_used = S->alloc.used();
    st->pool_peak = S->pool_peak;
    return VG_OK;
}

}  // extern "C"
