// Integrators and the Machine: the C++ counterpart of src/integrator.rs and
// src/machine.rs of the reference, with the sweeps on the GPU.
//
// Machine<H, I, S> (machine.rs:44-125) owns the thermostat, the Hamiltonian,
// the integrator kind, the instruments and the state; the state lives on the
// device as one chain of the engine.  relax_for / measure_for fire the same
// hooks in the same order; a call of run(rng, steps) becomes one engine call
// for all `steps` sweeps (split only where a StateSensor wants spins), and
// the per-sweep energy and |M| are fed to after_step.
//
// The engine's randomness is counter-based (Philox keyed by a 64-bit seed,
// sweep, site): the seed is drawn from the caller's rng at the first
// relax_for / measure_for; later calls continue the chain's counters.
#pragma once

#include <map>
#include <memory>
#include <random>
#include <vector>

#include "energy.hpp"
#include "instrument.hpp"

namespace vegas {

// Integrator kinds (integrator.rs:40-230).  The sweep itself runs in the
// engine; step() is the per-call adapter of Integrator::step for callers
// that hold a host State (it uploads, sweeps once, downloads).
struct MetropolisIntegrator {
    static constexpr int proposal = VG_PROPOSAL_RANDOM;
};
struct MetropolisFlipIntegrator {
    static constexpr int proposal = VG_PROPOSAL_FLIP;
};
// One "step" of a Wolff machine is one cluster grown and flipped
// (integrator.rs:172-230); Ising spins only.
struct WolffIntegrator {
    static constexpr int proposal = VG_PROPOSAL_WOLFF;
};

// A process-wide engine context on device 0.
inline vg_ctx* default_context() {
    static vg_ctx* ctx = [] {
        vg_ctx* c = nullptr;
        check(vg_ctx_create(0, &c));
        return c;
    }();
    return ctx;
}

// One chain of the engine: descriptor, device state, thermostat, runs.
template <class S>
class Chain {
  public:
    // matrix: the Exchange's coupling matrix when it was built from one
    // (the chain is then a graph chain; lattice is ignored).
    Chain(const vg_lattice& lattice, int proposal, const vg_hamiltonian& h, uint64_t seed,
          vg_ctx* ctx = default_context(), const CsrMatrix* matrix = nullptr) {
        vg_chain_desc d{};
        d.lattice = lattice;
        d.spin = S::kind;
        d.proposal = proposal;
        d.hamiltonian = h;
        d.seed = seed;
        if (matrix) {
            d.lattice = vg_lattice{VG_LATTICE_GRAPH, 0, 0, 0, 0};
            const vg_csr m = matrix->view();
            check(vg_batch_create_csr(ctx, 1, &d, &m, &batch_));
        } else {
            check(vg_batch_create(ctx, 1, &d, &batch_));
        }
        uint64_t n = 0;
        check(vg_batch_num_sites(batch_, 0, &n));
        n_ = n;
    }
    ~Chain() { vg_batch_destroy(batch_); }
    Chain(const Chain&) = delete;
    Chain& operator=(const Chain&) = delete;

    size_t num_sites() const { return n_; }
    void set_thermostat(const Thermostat<S>& th) {
        double dir[3];
        vector_of(th.field().orientation(), dir);
        check(vg_batch_set_thermostat(batch_, 0, th.temperature(), th.field().magnitude(), dir));
    }
    void upload(const State<S>& st) {
        if (st.len() != n_)
            throw Error(ErrorKind::InvalidArgument, "state size does not match the lattice");
        upload_impl(st);
    }
    State<S> download() const;
    void run(size_t sweeps, std::vector<SweepRecord>& out) {
        std::vector<double> e(sweeps), m(sweeps);
        check(vg_batch_run(batch_, sweeps, e.data(), m.data()));
        out.resize(sweeps);
        for (size_t k = 0; k < sweeps; ++k) out[k] = SweepRecord{e[k], m[k]};
    }
    vg_batch* handle() { return batch_; }
    // Sweeps run so far: the chain's Philox counter.
    uint64_t sweep_count() const {
        uint64_t n = 0;
        check(vg_batch_sweep_count(batch_, 0, &n));
        return n;
    }
    void set_sweep_count(uint64_t n) { check(vg_batch_set_sweep_count(batch_, 0, n)); }

  private:
    void upload_impl(const State<S>& st);
    vg_batch* batch_ = nullptr;
    size_t n_ = 0;
};

template <>
inline void Chain<IsingSpin>::upload_impl(const State<IsingSpin>& st) {
    std::vector<int8_t> v(n_);
    for (size_t i = 0; i < n_; ++i) v[i] = (int8_t)st.at(i).value();
    check(vg_batch_set_ising(batch_, 0, v.data()));
}
template <>
inline void Chain<HeisenbergSpin>::upload_impl(const State<HeisenbergSpin>& st) {
    std::vector<double> v(3 * n_);
    for (size_t i = 0; i < n_; ++i) {
        v[3 * i] = st.at(i).sx();
        v[3 * i + 1] = st.at(i).sy();
        v[3 * i + 2] = st.at(i).sz();
    }
    check(vg_batch_set_heisenberg(batch_, 0, v.data()));
}
template <>
inline State<IsingSpin> Chain<IsingSpin>::download() const {
    std::vector<int8_t> v(n_);
    check(vg_batch_get_ising(batch_, 0, v.data()));
    std::vector<IsingSpin> s;
    s.reserve(n_);
    for (int8_t x : v) s.push_back(IsingSpin(x > 0 ? IsingSpin::Up : IsingSpin::Down));
    return State<IsingSpin>(std::move(s));
}
template <>
inline State<HeisenbergSpin> Chain<HeisenbergSpin>::download() const {
    std::vector<double> v(3 * n_);
    check(vg_batch_get_heisenberg(batch_, 0, v.data()));
    std::vector<HeisenbergSpin> s;
    s.reserve(n_);
    for (size_t i = 0; i < n_; ++i) s.emplace_back(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    return State<HeisenbergSpin>(std::move(s));
}

// The lattice a Hamiltonian's Exchange term was built on; without an
// Exchange term the sites have no bonds and any layout of n sites will do.
template <class H>
vg_lattice lattice_of(const H& h, size_t n) {
    if (const Lattice* l = h.lattice()) return l->descriptor();
    if (h.matrix()) return vg_lattice{VG_LATTICE_GRAPH, 0, 0, 0, 0};
    if (n % 2 != 0)
        throw Error(ErrorKind::UnsupportedLattice,
                    "a Hamiltonian without Exchange needs an even number of sites");
    vg_lattice d{};
    d.kind = VG_LATTICE_SC;
    d.lx = (uint32_t)n;
    d.ly = d.lz = 1;
    d.pbc = 0;
    return d;
}

// Integrator::step (integrator.rs:40-49) for a host State: one sweep.
template <class I, class Rng, class H, class S>
State<S> step(const I&, Rng& rng, const Thermostat<S>& th, const H& h, const State<S>& st) {
    Chain<S> c(lattice_of(h, st.len()), I::proposal, describe(h), rng(), default_context(),
               h.matrix());
    c.upload(st);
    c.set_thermostat(th);
    std::vector<SweepRecord> r;
    c.run(1, r);
    return c.download();
}

// A Machine's chain between two stages (checkpoint.hpp saves and loads
// these): the chain's Philox key and counter, its spins and thermostat, and
// the index of the next stage of the program being run.
template <class S>
struct Checkpoint {
    bool started = false;  // the chain exists (seed and sweeps are meaningful)
    uint64_t seed = 0;
    uint64_t sweeps = 0;
    uint64_t next_stage = 0;
    Thermostat<S> thermostat = Thermostat<S>::near_zero();
    State<S> state;
};

template <class H, class I, class S>
class Machine {
  public:
    using Instruments = std::vector<std::unique_ptr<Instrument<H, S>>>;
    Machine(Thermostat<S> thermostat, H hamiltonian, I integrator, Instruments instruments,
            State<S> state)
        : thermostat_(thermostat),
          hamiltonian_(std::move(hamiltonian)),
          integrator_(integrator),
          instruments_(std::move(instruments)),
          state_(std::move(state)) {}

    const Thermostat<S>& thermostat() const { return thermostat_; }
    void set_thermostat(Thermostat<S> th) { thermostat_ = th; }
    const H& hamiltonian() const { return hamiltonian_; }
    // The current spins (downloaded from the device once sweeps have run).
    State<S> state() const { return chain_ ? chain_->download() : state_; }

    template <class Rng>
    void relax_for(Rng& rng, size_t steps) {
        for (auto& i : instruments_) i->on_relax_start(thermostat_, hamiltonian_, state_.len());
        run(rng, steps);
        for (auto& i : instruments_) i->on_relax_end();
    }
    template <class Rng>
    void measure_for(Rng& rng, size_t steps) {
        for (auto& i : instruments_) i->on_measure_start(thermostat_, hamiltonian_, state_.len());
        run(rng, steps);
        for (auto& i : instruments_) i->on_measure_end();
    }
    Chain<S>& chain() { return *chain_; }

    // The machine's chain now, with next_stage as the program cursor.
    Checkpoint<S> checkpoint(uint64_t next_stage) const {
        Checkpoint<S> ck;
        ck.next_stage = next_stage;
        ck.thermostat = thermostat_;
        if (chain_) {
            ck.started = true;
            ck.seed = seed_;
            ck.sweeps = chain_->sweep_count();
            ck.state = chain_->download();
        } else {
            ck.state = state_;
        }
        return ck;
    }
    // Continues from a checkpoint: the chain gets the checkpoint's seed,
    // counter and spins (no seed is drawn from the caller's rng), so its
    // next sweeps are the ones the checkpointed run would have made.
    void restore(const Checkpoint<S>& ck) {
        if (ck.state.len() != state_.len())
            throw Error(ErrorKind::InvalidArgument, "checkpoint size does not match the machine");
        thermostat_ = ck.thermostat;
        state_ = ck.state;
        chain_.reset();
        if (ck.started) {
            make_chain(ck.seed);
            chain_->set_sweep_count(ck.sweeps);
        }
    }

  private:
    void make_chain(uint64_t seed) {
        seed_ = seed;
        chain_ = std::make_unique<Chain<S>>(lattice_of(hamiltonian_, state_.len()), I::proposal,
                                            describe(hamiltonian_), seed, default_context(),
                                            hamiltonian_.matrix());
        chain_->upload(state_);
    }

    template <class Rng>
    void run(Rng& rng, size_t steps) {
        if (!chain_) make_chain((uint64_t)rng());
        chain_->set_thermostat(thermostat_);
        std::vector<SweepRecord> rec;
        size_t done = 0;
        while (done < steps) {
            // Stop right after the next sweep whose phase step some
            // instrument wants spins for (StateSensor frequency).
            size_t n = steps - done;
            for (auto& i : instruments_) {
                const size_t f = i->state_frequency();
                if (f) n = std::min(n, (done + f - 1) / f * f - done + 1);
            }
            chain_->run(n, rec);
            for (size_t k = 0; k < n; ++k) {
                const size_t phase_step = done + k;
                if (k + 1 == n) {
                    bool any = false;
                    for (auto& i : instruments_) {
                        const size_t f = i->state_frequency();
                        if (f && phase_step % f == 0) any = true;
                    }
                    if (any) {
                        const State<S> st = chain_->download();
                        for (auto& i : instruments_) {
                            const size_t f = i->state_frequency();
                            if (f && phase_step % f == 0) i->on_state(st);
                        }
                    }
                }
                for (auto& i : instruments_) i->after_step(rec[k]);
            }
            done += n;
        }
    }

    Thermostat<S> thermostat_;
    H hamiltonian_;
    I integrator_;
    Instruments instruments_;
    State<S> state_;
    std::unique_ptr<Chain<S>> chain_;
    uint64_t seed_ = 0;  // the chain's Philox key
};

}  // namespace vegas
