// Streaming moments (reference src/accumulator.rs:23-63): sum, sum of
// squares, sum of fourth powers, count; variance = <x^2> - <x>^2 and the
// Binder cumulant 1 - <x^4> / (3 <x^2>^2) from the naive one-pass sums.
#pragma once

#include <cstddef>

namespace vegas {

class Accumulator {
  public:
    void collect(double v) {
        sum_ += v;
        sum_sq_ += v * v;
        sum_fourth_ += v * v * v * v;
        ++count_;
    }
    double mean() const { return sum_ / (double)count_; }
    double variance() const {
        const double m = mean();
        return sum_sq_ / (double)count_ - m * m;
    }
    double binder_cumulant() const {
        const double m2 = sum_sq_ / (double)count_;
        return 1.0 - (sum_fourth_ / (double)count_) / (3.0 * (m2 * m2));
    }
    size_t count() const { return count_; }

  private:
    double sum_ = 0.0, sum_sq_ = 0.0, sum_fourth_ = 0.0;
    size_t count_ = 0;
};

}  // namespace vegas
