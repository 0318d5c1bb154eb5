// Instruments hooked into a Machine (reference src/instrument.rs:19-351).
//
// The hooks fire in the reference's order (machine.rs:91-125): on_*_start,
// then after_step once per sweep, then on_*_end.  Where the reference hands
// each hook the whole State and lets every instrument recompute
// H.total_energy and the magnetization (instrument.rs:133-141, 254-262), the
// engine computes both on the device per sweep, by the same conventions, and
// hands them over as a SweepRecord.  An instrument that wants spins asks for
// them with state_frequency() and receives on_state() on those sweeps.
#pragma once

#include <cmath>
#include <cstdio>
#include <ostream>
#include <string>
#include <vector>

#include <memory>

#include "accumulator.hpp"
#include "parquet.hpp"
#include "thermostat.hpp"

namespace vegas {

struct SweepRecord {
    double energy;         // H.total_energy of the state after the sweep
    double magnetization;  // state.magnetization().magnitude()
};

template <class H, class S>
class Instrument {
  public:
    virtual ~Instrument() = default;
    virtual void on_relax_start(const Thermostat<S>&, const H&, size_t /*n*/) {}
    virtual void on_relax_end() {}
    virtual void on_measure_start(const Thermostat<S>&, const H&, size_t /*n*/) {}
    virtual void on_measure_end() {}
    virtual void after_step(const SweepRecord&) {}
    // > 0: on_state() is called after every sweep whose phase step is a
    // multiple of it (StateSensor's frequency, instrument.rs:340-350).
    virtual size_t state_frequency() const { return 0; }
    virtual void on_state(const State<S>&) {}
};

// Rust's `{:.16}` for an f64: fixed notation, 16 decimals, NaN and inf as
// Rust spells them.
inline std::string format16(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.16f", v);
    return buf;
}

// StatSensor (instrument.rs:61-142): during measurement, accumulates E and
// |M| per sweep and at the end writes "T |h| <E> C <|M|> chi U4" with
// C = Var(E)/(N T^2), chi = Var(|M|)/(N T), U4 the Binder cumulant of |M|.
template <class H, class S>
class StatSensor : public Instrument<H, S> {
  public:
    explicit StatSensor(std::ostream& out) : out_(out) {}
    void on_measure_start(const Thermostat<S>& th, const H&, size_t n) override {
        active_ = true;
        temperature_ = th.temperature();
        field_ = th.field().magnitude();
        n_ = n;
    }
    void on_measure_end() override {
        if (active_) {
            const double n = (double)n_, T = temperature_;
            out_ << format16(T) << ' ' << format16(field_) << ' ' << format16(energy_.mean())
                 << ' ' << format16(energy_.variance() / (n * (T * T))) << ' '
                 << format16(magnetization_.mean()) << ' '
                 << format16(magnetization_.variance() / (n * T)) << ' '
                 << format16(magnetization_.binder_cumulant()) << '\n';
        }
        active_ = false;
        energy_ = Accumulator();
        magnetization_ = Accumulator();
    }
    void after_step(const SweepRecord& r) override {
        if (!active_) return;
        energy_.collect(r.energy);
        magnetization_.collect(r.magnetization);
    }

  private:
    std::ostream& out_;
    bool active_ = false;
    double temperature_ = 0.0, field_ = 0.0;
    size_t n_ = 0;
    Accumulator energy_, magnetization_;
};

// One row of the observables table (the parquet schema of output.rs:37-46).
struct ObservableRow {
    bool relax;
    uint64_t stage, step, n;
    double temperature, field, energy, magnetization;
};

// ObservableSensor (instrument.rs:145-263): per sweep of every relax and
// measure phase, a row; the stage counter grows by one per phase.  Built
// with a path, it writes each phase's rows to that parquet file at the
// phase's end (instrument.rs:196-205, 227-236) and finishes the file when
// closed or destroyed; built without, it keeps every row in memory.
template <class H, class S>
class ObservableSensor : public Instrument<H, S> {
  public:
    ObservableSensor() = default;
    explicit ObservableSensor(const std::string& parquet_path)
        : io_(std::make_unique<ParquetWriter>(parquet_path, observable_schema())) {}
    void on_relax_start(const Thermostat<S>& th, const H&, size_t n) override { begin(th, n, true); }
    void on_relax_end() override { end(); }
    void on_measure_start(const Thermostat<S>& th, const H&, size_t n) override {
        begin(th, n, false);
    }
    void on_measure_end() override { end(); }
    void after_step(const SweepRecord& r) override {
        if (!active_) return;
        rows_.push_back(ObservableRow{relax_, stage_, step_++, n_, temperature_, field_, r.energy,
                                      r.magnetization});
    }
    // Every row (no path), or the rows of the phase in progress (path).
    const std::vector<ObservableRow>& rows() const { return rows_; }
    // Finishes the parquet file (no-op without a path).
    void close() {
        if (io_) io_->close();
    }
    // Writes the rows as CSV (header = the parquet column names) to
    // `path`.tmp and renames it into place, as the parquet writer does on
    // close (output.rs:101-113).
    void write_csv(const std::string& path) const {
        const std::string tmp = path + ".tmp";
        FILE* f = std::fopen(tmp.c_str(), "w");
        if (!f) throw Error(ErrorKind::Io, "cannot open " + tmp);
        std::fprintf(f, "relax,stage,step,n,temperature,field,energy,magnetization\n");
        for (const auto& r : rows_)
            std::fprintf(f, "%s,%llu,%llu,%llu,%.17g,%.17g,%.17g,%.17g\n", r.relax ? "true" : "false",
                         (unsigned long long)r.stage, (unsigned long long)r.step,
                         (unsigned long long)r.n, r.temperature, r.field, r.energy, r.magnetization);
        std::fclose(f);
        if (std::rename(tmp.c_str(), path.c_str()) != 0)
            throw Error(ErrorKind::Io, "cannot rename " + tmp);
    }

  private:
    void begin(const Thermostat<S>& th, size_t n, bool relax) {
        active_ = true;
        relax_ = relax;
        temperature_ = th.temperature();
        field_ = th.field().magnitude();
        n_ = n;
        step_ = 0;
    }
    void end() {
        if (io_ && !rows_.empty()) {
            // one row group per phase, as the reference's write() per phase
            const size_t m = rows_.size();
            std::vector<uint8_t> relax(m);
            std::vector<uint64_t> stage(m), step(m), n(m);
            std::vector<double> t(m), h(m), e(m), mag(m);
            for (size_t i = 0; i < m; ++i) {
                const ObservableRow& r = rows_[i];
                relax[i] = r.relax;
                stage[i] = r.stage;
                step[i] = r.step;
                n[i] = r.n;
                t[i] = r.temperature;
                h[i] = r.field;
                e[i] = r.energy;
                mag[i] = r.magnetization;
            }
            io_->write(m, {relax.data(), stage.data(), step.data(), n.data(), t.data(), h.data(),
                           e.data(), mag.data()});
            rows_.clear();
        }
        active_ = false;
        ++stage_;
    }
    std::unique_ptr<ParquetWriter> io_;
    std::vector<ObservableRow> rows_;
    bool active_ = false, relax_ = false;
    uint64_t stage_ = 0, step_ = 0, n_ = 0;
    double temperature_ = 0.0, field_ = 0.0;
};

// StateSensor (instrument.rs:265-351): the spins every `frequency` sweeps of
// each phase.  Built with a path, each snapshot becomes one row per site of
// that parquet file (relax, stage, step, temperature, field, id, sx, sy, sz;
// output.rs:155-166); built without, snapshots are kept in memory.
template <class H, class S>
class StateSensor : public Instrument<H, S> {
  public:
    struct Snapshot {
        bool relax;
        uint64_t stage, step;
        double temperature, field;
        State<S> state;
    };
    explicit StateSensor(size_t frequency) : frequency_(frequency) {
        if (frequency == 0) throw Error(ErrorKind::InvalidArgument, "frequency must be positive");
    }
    StateSensor(const std::string& parquet_path, size_t frequency) : StateSensor(frequency) {
        io_ = std::make_unique<ParquetWriter>(parquet_path, state_schema());
    }
    void on_relax_start(const Thermostat<S>& th, const H&, size_t) override { begin(th, true); }
    void on_relax_end() override { end(); }
    void on_measure_start(const Thermostat<S>& th, const H&, size_t) override { begin(th, false); }
    void on_measure_end() override { end(); }
    void after_step(const SweepRecord&) override { ++step_; }
    size_t state_frequency() const override { return frequency_; }
    void on_state(const State<S>& st) override {
        if (!active_) return;
        if (io_)
            write(st);
        else
            snapshots_.push_back(Snapshot{relax_, stage_, step_, temperature_, field_, st});
    }
    const std::vector<Snapshot>& snapshots() const { return snapshots_; }
    void close() {
        if (io_) io_->close();
    }

  private:
    void write(const State<S>& st) {
        const size_t m = st.len();
        std::vector<uint8_t> relax(m, relax_ ? 1 : 0);
        std::vector<uint64_t> stage(m, stage_), step(m, step_), id(m);
        std::vector<double> t(m, temperature_), h(m, field_), sx(m), sy(m), sz(m);
        for (size_t i = 0; i < m; ++i) {
            const S& s = st.at(i);
            id[i] = i;
            sx[i] = s.sx();
            sy[i] = s.sy();
            sz[i] = s.sz();
        }
        io_->write(m, {relax.data(), stage.data(), step.data(), t.data(), h.data(), id.data(),
                       sx.data(), sy.data(), sz.data()});
    }
    void begin(const Thermostat<S>& th, bool relax) {
        active_ = true;
        relax_ = relax;
        temperature_ = th.temperature();
        field_ = th.field().magnitude();
        step_ = 0;
    }
    void end() {
        active_ = false;
        step_ = 0;
        ++stage_;
    }
    size_t frequency_;
    std::unique_ptr<ParquetWriter> io_;
    std::vector<Snapshot> snapshots_;
    bool active_ = false, relax_ = false;
    uint64_t stage_ = 0, step_ = 0;
    double temperature_ = 0.0, field_ = 0.0;
};

}  // namespace vegas
