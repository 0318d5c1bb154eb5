// Spins, fields and states: the C++ counterpart of src/state.rs and
// src/util.rs of the reference.  IsingSpin (state.rs:61-130) and
// HeisenbergSpin (state.rs:134-192) keep the reference's algebra: Ising dot
// products are +-1, Field::magnitude() is |magnitude| (state.rs:219-221), the
// magnetization is the vector sum passed through from_projections
// (state.rs:86-92, 150-160, 235-242, 291-296).
#pragma once

#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "vegas_hip.h"

namespace vegas {

// Marsaglia (1972): uniform point on S^2 by rejection (util.rs:21-34).
template <class Rng>
std::array<double, 3> marsaglia(Rng& rng) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (;;) {
        const double x1 = u(rng), x2 = u(rng);
        if (x1 * x1 + x2 * x2 >= 1.0) continue;
        const double r = std::sqrt(1.0 - x1 * x1 - x2 * x2);
        return {2.0 * x1 * r, 2.0 * x2 * r, 1.0 - 2.0 * (x1 * x1 + x2 * x2)};
    }
}

template <class S>
class Field;

class IsingSpin {
  public:
    enum Value : int8_t { Up = 1, Down = -1 };
    static constexpr int kind = VG_SPIN_ISING;
    IsingSpin(Value v = Up) : v_(v) {}
    static IsingSpin up() { return IsingSpin(Up); }
    static IsingSpin down() { return IsingSpin(Down); }
    // Up when a uniform [0,1) draw is below 1/2 (state.rs:76-84).
    template <class Rng>
    static IsingSpin rand(Rng& rng) {
        std::uniform_real_distribution<double> u(0.0, 1.0);
        return IsingSpin(u(rng) < 0.5 ? Up : Down);
    }
    static Field<IsingSpin> from_projections(double sx, double sy, double sz);
    IsingSpin flip() const { return IsingSpin(v_ == Up ? Down : Up); }
    double dot(const IsingSpin& o) const { return v_ == o.v_ ? 1.0 : -1.0; }
    double sx() const { return 0.0; }
    double sy() const { return 0.0; }
    double sz() const { return v_ == Up ? 1.0 : -1.0; }
    Value value() const { return v_; }
    bool operator==(const IsingSpin& o) const { return v_ == o.v_; }
    bool operator!=(const IsingSpin& o) const { return v_ != o.v_; }

  private:
    Value v_;
};

class HeisenbergSpin {
  public:
    static constexpr int kind = VG_SPIN_HEISENBERG;
    HeisenbergSpin(double x = 0.0, double y = 0.0, double z = 1.0) : s_{x, y, z} {}
    static HeisenbergSpin up() { return HeisenbergSpin(0.0, 0.0, 1.0); }
    static HeisenbergSpin down() { return HeisenbergSpin(0.0, 0.0, -1.0); }
    template <class Rng>
    static HeisenbergSpin rand(Rng& rng) {
        const auto v = marsaglia(rng);
        return HeisenbergSpin(v[0], v[1], v[2]);
    }
    static Field<HeisenbergSpin> from_projections(double sx, double sy, double sz);
    HeisenbergSpin flip() const { return HeisenbergSpin(-s_[0], -s_[1], -s_[2]); }
    double dot(const HeisenbergSpin& o) const {
        return s_[0] * o.s_[0] + s_[1] * o.s_[1] + s_[2] * o.s_[2];
    }
    double sx() const { return s_[0]; }
    double sy() const { return s_[1]; }
    double sz() const { return s_[2]; }
    bool operator==(const HeisenbergSpin& o) const { return s_ == o.s_; }
    bool operator!=(const HeisenbergSpin& o) const { return s_ != o.s_; }

  private:
    std::array<double, 3> s_;
};

// Orientation plus magnitude; magnitude() reads |magnitude| (state.rs:219-221).
template <class S>
class Field {
  public:
    Field() : orientation_(S::up()), magnitude_(0.0) {}
    Field(S orientation, double magnitude) : orientation_(orientation), magnitude_(magnitude) {}
    static Field zero() { return Field(); }
    double magnitude() const { return std::fabs(magnitude_); }
    const S& orientation() const { return orientation_; }

  private:
    S orientation_;
    double magnitude_;
};

inline Field<IsingSpin> IsingSpin::from_projections(double, double, double sz) {
    return Field<IsingSpin>(sz >= 0.0 ? IsingSpin(Up) : IsingSpin(Down), std::fabs(sz));
}

inline Field<HeisenbergSpin> HeisenbergSpin::from_projections(double sx, double sy, double sz) {
    const double m = std::sqrt(sx * sx + sy * sy + sz * sz);
    if (std::fabs(m) < DBL_EPSILON) return Field<HeisenbergSpin>::zero();
    return Field<HeisenbergSpin>(HeisenbergSpin(sx / m, sy / m, sz / m), m);
}

template <class S>
class State {
  public:
    State() = default;
    explicit State(std::vector<S> spins) : spins_(std::move(spins)) {}
    static State up_with_size(size_t n) { return State(std::vector<S>(n, S::up())); }
    static State down_with_size(size_t n) { return State(std::vector<S>(n, S::down())); }
    template <class Rng>
    static State rand_with_size(Rng& rng, size_t n) {
        std::vector<S> v;
        v.reserve(n);
        for (size_t i = 0; i < n; ++i) v.push_back(S::rand(rng));
        return State(std::move(v));
    }
    const std::vector<S>& spins() const { return spins_; }
    const S& at(size_t i) const { return spins_.at(i); }
    void set_at(size_t i, S s) { spins_.at(i) = s; }
    size_t len() const { return spins_.size(); }
    bool is_empty() const { return spins_.empty(); }
    Field<S> magnetization() const {
        double x = 0.0, y = 0.0, z = 0.0;
        for (const S& s : spins_) {
            x += s.sx();
            y += s.sy();
            z += s.sz();
        }
        return S::from_projections(x, y, z);
    }

  private:
    std::vector<S> spins_;
};

}  // namespace vegas
