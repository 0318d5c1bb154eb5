// Thermal bath: temperature clamped to at least DBL_EPSILON and a field
// (reference src/thermostat.rs:19-78).
#pragma once

#include <cfloat>

#include "state.hpp"

namespace vegas {

template <class S>
class Thermostat {
  public:
    Thermostat(double temperature, Field<S> field)
        : temperature_(temperature < DBL_EPSILON ? DBL_EPSILON : temperature), field_(field) {}
    static Thermostat near_zero() { return Thermostat(DBL_EPSILON, Field<S>::zero()); }
    Thermostat with_temperature(double temperature) const { return Thermostat(temperature, field_); }
    Thermostat with_field(Field<S> field) const {
        Thermostat t = *this;
        t.field_ = field;
        return t;
    }
    double temperature() const { return temperature_; }
    const Field<S>& field() const { return field_; }

  private:
    double temperature_;
    Field<S> field_;
};

}  // namespace vegas
