// Hamiltonian terms: the C++ counterpart of src/energy.rs of the reference.
//
// Each term evaluates the reference's per-site energy on a host State (the
// API of Hamiltonian::energy / total_energy, energy.rs:45-60) and describes
// itself to the engine (vg_hamiltonian) so the GPU sweeps use the same local
// energy.  The reference's conventions are kept, quirks included:
//   Zeeman per-site energy is +|h| s.h while its total_energy is -|h| sum s.h
//   (energy.rs:147-160); Anisotropy's total_energy omits K (energy.rs:114-120);
//   Exchange halves its site sum (energy.rs:208-213); Compound uses the
//   default sum of site energies (energy.rs:55-59, 248-257).
#pragma once

#include <memory>
#include <utility>

#include "lattice.hpp"
#include "matrix.hpp"
#include "thermostat.hpp"

namespace vegas {

// Default Hamiltonian::total_energy: sum of site energies (energy.rs:55-59).
template <class H, class S>
double site_sum(const H& h, const Thermostat<S>& th, const State<S>& st) {
    double e = 0.0;
    for (size_t i = 0; i < st.len(); ++i) e += h.energy(th, st, i);
    return e;
}

inline void vector_of(const IsingSpin& s, double v[3]) {
    v[0] = 0.0;
    v[1] = 0.0;
    v[2] = s.sz();
}
inline void vector_of(const HeisenbergSpin& s, double v[3]) {
    v[0] = s.sx();
    v[1] = s.sy();
    v[2] = s.sz();
}

// Gauge (energy.rs:64-80): constant site energy.
class Gauge {
  public:
    explicit Gauge(double value) : value_(value) {}
    template <class S>
    double energy(const Thermostat<S>&, const State<S>&, size_t) const { return value_; }
    template <class S>
    double total_energy(const Thermostat<S>& th, const State<S>& st) const {
        return site_sum(*this, th, st);
    }
    void describe(vg_hamiltonian& d) const {
        d.terms |= VG_TERM_GAUGE;
        d.gauge = value_;
    }
    const Lattice* lattice() const { return nullptr; }
    const CsrMatrix* matrix() const { return nullptr; }

  private:
    double value_;
};

// UniaxialAnisotropy (energy.rs:84-121): K (s.reference)^2; total omits K.
template <class S>
class UniaxialAnisotropy {
  public:
    UniaxialAnisotropy(S reference, double k) : reference_(reference), k_(k) {}
    double energy(const Thermostat<S>&, const State<S>& st, size_t i) const {
        const double d = st.at(i).dot(reference_);
        return d * d * k_;
    }
    double total_energy(const Thermostat<S>&, const State<S>& st) const {
        double e = 0.0;
        for (const S& s : st.spins()) {
            const double d = s.dot(reference_);
            e += d * d;
        }
        return e;
    }
    void describe(vg_hamiltonian& d) const {
        d.terms |= VG_TERM_ANISOTROPY;
        d.anisotropy = k_;
        vector_of(reference_, d.anisotropy_axis);
    }
    const Lattice* lattice() const { return nullptr; }
    const CsrMatrix* matrix() const { return nullptr; }

  private:
    S reference_;
    double k_;
};

// Zeeman (energy.rs:125-161): site energy +|h| s.h, total -|h| sum s.h.
template <class S>
class Zeeman {
  public:
    double energy(const Thermostat<S>& th, const State<S>& st, size_t i) const {
        return st.at(i).dot(th.field().orientation()) * th.field().magnitude();
    }
    double total_energy(const Thermostat<S>& th, const State<S>& st) const {
        double e = 0.0;
        for (const S& s : st.spins()) e += s.dot(th.field().orientation());
        return -th.field().magnitude() * e;
    }
    void describe(vg_hamiltonian& d) const { d.terms |= VG_TERM_ZEEMAN; }
    const Lattice* lattice() const { return nullptr; }
    const CsrMatrix* matrix() const { return nullptr; }
};

// Exchange (energy.rs:165-214): e_i = -sum_j J_ij s_i.s_j over row i of the
// coupling matrix; total = sum_i e_i / 2.  Built from a lattice (every
// neighbour slot with J, duplicates counted, as the CSR built by
// Exchange::from_lattice sums them; the engine sweeps it by its sc/bcc/fcc
// stencil) or from any matrix (Exchange::new, energy.rs:170-172; the engine
// sweeps it as a graph chain, by colour classes).
class Exchange {
  public:
    static Exchange from_lattice(double exchange, const Lattice& lattice) {
        return Exchange(exchange, lattice);
    }
    // Exchange::from_lattice with an explicit edge filter (matrix.hpp):
    // None is from_lattice above; SourceNotAfterTarget applies the
    // reference's `source <= target` filter to the unit cell's edges and runs
    // the resulting matrix.
    static Exchange from_lattice(double exchange, const Lattice& lattice, EdgeFilter filter) {
        if (filter == EdgeFilter::None) return Exchange(exchange, lattice);
        return from_matrix(lattice_matrix(lattice, exchange, filter));
    }
    // Exchange::new(CsMat<f64>) (energy.rs:170-172): the rows as given.
    static Exchange from_matrix(CsrMatrix matrix) {
        Exchange e(1.0, Lattice::sc());
        e.matrix_ = std::make_shared<const CsrMatrix>(std::move(matrix));
        return e;
    }
    template <class S>
    double energy(const Thermostat<S>&, const State<S>& st, size_t i) const {
        double e = 0.0;
        if (matrix_) {
            for (uint64_t k = matrix_->row_ptr[i]; k < matrix_->row_ptr[i + 1]; ++k)
                e += -matrix_->value[k] * st.at(i).dot(st.at((size_t)matrix_->col[k]));
            return e;
        }
        for (int64_t j : lattice_.neighbours(i))
            if (j >= 0) e += -j_ * st.at(i).dot(st.at((size_t)j));
        return e;
    }
    template <class S>
    double total_energy(const Thermostat<S>& th, const State<S>& st) const {
        return site_sum(*this, th, st) / 2.0;
    }
    void describe(vg_hamiltonian& d) const {
        d.terms |= VG_TERM_EXCHANGE;
        d.exchange = matrix_ ? 1.0 : j_;  // a matrix carries its own couplings
    }
    const Lattice* lattice() const { return matrix_ ? nullptr : &lattice_; }
    const CsrMatrix* matrix() const { return matrix_.get(); }

  private:
    Exchange(double j, Lattice l) : j_(j), lattice_(l) {}
    double j_;
    Lattice lattice_;
    std::shared_ptr<const CsrMatrix> matrix_;
};

// Compound (energy.rs:221-257): e_i = a_i + b_i, default total.
template <class A, class B>
class Compound {
  public:
    Compound(A a, B b) : a_(std::move(a)), b_(std::move(b)) {}
    template <class S>
    double energy(const Thermostat<S>& th, const State<S>& st, size_t i) const {
        return a_.energy(th, st, i) + b_.energy(th, st, i);
    }
    template <class S>
    double total_energy(const Thermostat<S>& th, const State<S>& st) const {
        return site_sum(*this, th, st);
    }
    void describe(vg_hamiltonian& d) const {
        a_.describe(d);
        b_.describe(d);
        d.compound = 1;
    }
    const Lattice* lattice() const { return a_.lattice() ? a_.lattice() : b_.lattice(); }
    const CsrMatrix* matrix() const { return a_.matrix() ? a_.matrix() : b_.matrix(); }

  private:
    A a_;
    B b_;
};

// hamiltonian!(a, b, c, ...) (energy.rs:273-290): left-nested Compounds.
template <class A>
A hamiltonian(A a) {
    return a;
}
template <class A, class B, class... R>
auto hamiltonian(A a, B b, R... rest) {
    return hamiltonian(Compound<A, B>(std::move(a), std::move(b)), std::move(rest)...);
}

// The engine's description of a Hamiltonian value.
template <class H>
vg_hamiltonian describe(const H& h) {
    vg_hamiltonian d{};
    h.describe(d);
    return d;
}

}  // namespace vegas
