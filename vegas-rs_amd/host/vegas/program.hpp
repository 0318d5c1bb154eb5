// Programs (reference src/program.rs:55-336): Relax, CoolDown and
// HysteresisLoop with the reference's defaults, setters, validation order
// (ProgramError kinds) and f64 schedules -- loops that step the temperature
// or field and test after stepping, so e.g. CoolDown 4.0 -> 0.1 by 0.1 runs
// 39 temperatures and HysteresisLoop's field overshoots max_field once.
// schedule() lists the stages run() will execute, for callers that drive the
// engine themselves (the job server).
#pragma once

#include <cfloat>
#include <vector>

#include "machine.hpp"

namespace vegas {

// One phase pair of a program: set the thermostat, relax, then measure
// (measure == 0 with `relax_only` for Relax).
struct Stage {
    double temperature;
    bool sets_field;     // the stage replaces the thermostat's field
    double field;        // Field::new(S::up(), field): signed, |field| is applied
    size_t relax;
    size_t measure;
    bool relax_only;     // Relax: one relax_for, no measure_for
};

class Relax {
  public:
    explicit Relax(size_t steps = 1000, double temperature = 3.0)
        : steps_(steps), temperature_(temperature) {}
    Relax set_steps(size_t v) const { Relax r = *this; r.steps_ = v; return r; }
    Relax set_temperature(double v) const { Relax r = *this; r.temperature_ = v; return r; }
    void validate() const {
        if (steps_ == 0) throw Error(ErrorKind::NoSteps);
        if (temperature_ < DBL_EPSILON) throw Error(ErrorKind::ZeroTemperature);
    }
    std::vector<Stage> schedule() const {
        validate();
        return {Stage{temperature_, false, 0.0, steps_, 0, true}};
    }
    template <class Rng, class H, class I, class S>
    void run(Rng& rng, Machine<H, I, S>& m) const {
        validate();
        m.set_thermostat(m.thermostat().with_temperature(temperature_));
        m.relax_for(rng, steps_);
    }

  private:
    size_t steps_;
    double temperature_;
};

class CoolDown {
  public:
    CoolDown(double max_temperature = 3.0, double min_temperature = 0.05, double cool_rate = 0.1,
             size_t relax = 1000, size_t steps = 20000)
        : max_(max_temperature), min_(min_temperature), rate_(cool_rate), relax_(relax),
          steps_(steps) {}
    CoolDown set_max_temperature(double v) const { CoolDown c = *this; c.max_ = v; return c; }
    CoolDown set_min_temperature(double v) const { CoolDown c = *this; c.min_ = v; return c; }
    CoolDown set_cool_rate(double v) const { CoolDown c = *this; c.rate_ = v; return c; }
    CoolDown set_relax(size_t v) const { CoolDown c = *this; c.relax_ = v; return c; }
    CoolDown set_steps(size_t v) const { CoolDown c = *this; c.steps_ = v; return c; }
    void validate() const {
        if (max_ < min_) throw Error(ErrorKind::TemperatureMaxLessThanMin);
        if (steps_ == 0) throw Error(ErrorKind::NoSteps);
        if (min_ < DBL_EPSILON) throw Error(ErrorKind::ZeroTemperature);
        if (rate_ < DBL_EPSILON) throw Error(ErrorKind::ZeroCoolRate);
    }
    std::vector<Stage> schedule() const {
        validate();
        std::vector<Stage> out;
        double t = max_;
        for (;;) {
            out.push_back(Stage{t, false, 0.0, relax_, steps_, false});
            t -= rate_;
            if (t < min_) break;
        }
        return out;
    }
    template <class Rng, class H, class I, class S>
    void run(Rng& rng, Machine<H, I, S>& m) const {
        for (const Stage& st : schedule()) {
            m.set_thermostat(m.thermostat().with_temperature(st.temperature));
            m.relax_for(rng, st.relax);
            m.measure_for(rng, st.measure);
        }
    }

  private:
    double max_, min_, rate_;
    size_t relax_, steps_;
};

class HysteresisLoop {
  public:
    HysteresisLoop(size_t steps = 1000, size_t relax = 1000, double temperature = 3.0,
                   double max_field = 1.0, double field_step = 0.1)
        : steps_(steps), relax_(relax), temperature_(temperature), max_field_(max_field),
          field_step_(field_step) {}
    HysteresisLoop set_steps(size_t v) const { HysteresisLoop h = *this; h.steps_ = v; return h; }
    HysteresisLoop set_relax(size_t v) const { HysteresisLoop h = *this; h.relax_ = v; return h; }
    HysteresisLoop set_temperature(double v) const {
        HysteresisLoop h = *this;
        h.temperature_ = v;
        return h;
    }
    HysteresisLoop set_max_field(double v) const {
        HysteresisLoop h = *this;
        h.max_field_ = v;
        return h;
    }
    HysteresisLoop set_field_step(double v) const {
        HysteresisLoop h = *this;
        h.field_step_ = v;
        return h;
    }
    void validate() const {
        if (steps_ == 0) throw Error(ErrorKind::NoSteps);
        if (temperature_ < DBL_EPSILON) throw Error(ErrorKind::ZeroTemperature);
        if (max_field_ < DBL_EPSILON) throw Error(ErrorKind::ZeroField);
        if (field_step_ < DBL_EPSILON) throw Error(ErrorKind::ZeroFieldStep);
    }
    std::vector<Stage> schedule() const {
        validate();
        std::vector<Stage> out;
        double h = 0.0;
        for (;;) {  // up to +max
            out.push_back(Stage{temperature_, true, h, relax_, steps_, false});
            h += field_step_;
            if (h > max_field_) break;
        }
        for (;;) {  // down to -max
            out.push_back(Stage{temperature_, true, h, relax_, steps_, false});
            h -= field_step_;
            if (h < -max_field_) break;
        }
        for (;;) {  // back up
            out.push_back(Stage{temperature_, true, h, relax_, steps_, false});
            h += field_step_;
            if (h > max_field_) break;
        }
        return out;
    }
    template <class Rng, class H, class I, class S>
    void run(Rng& rng, Machine<H, I, S>& m) const {
        const std::vector<Stage> stages = schedule();
        m.set_thermostat(m.thermostat().with_temperature(temperature_));
        for (const Stage& st : stages) {
            m.set_thermostat(m.thermostat().with_field(Field<S>(S::up(), st.field)));
            m.relax_for(rng, st.relax);
            m.measure_for(rng, st.measure);
        }
    }

  private:
    size_t steps_, relax_;
    double temperature_, max_field_, field_step_;
};

}  // namespace vegas
