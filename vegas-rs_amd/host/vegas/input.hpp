// Simulation input (reference src/input.rs): the TOML document that `vegas
// run` reads, its defaults (`vegas input` prints them), and the run itself
// (Input::run_with_spin, input.rs:264-294) over the GPU Machine.
//
// The document follows the reference's serde layout: `model` ("Ising" |
// "Heisenberg"), `algorithm` ("Metropolis" | "Wolff"), optional `exchange`,
// `[sample]` with `unitcell` ({name = "sc" | "bcc" | "fcc"} or {path}),
// `size` {x, y, z} and `pbc` {x, y, z}, `[[stages]]` tagged by `program`
// ("Relax" | "CoolDown" | "Hysteresis") with the program's fields, and an
// optional `[output]` with `observables` and `[output.state]` {path,
// frequency}.  As with serde, every non-optional field is required, unknown
// keys are ignored, and integers are accepted where a float is expected.
#pragma once

#include <fstream>
#include <iostream>
#include <memory>
#include <optional>
#include <random>
#include <sstream>
#include <string>
#include <variant>
#include <vector>

#include "instrument.hpp"
#include "machine.hpp"
#include "program.hpp"
#include "toml.hpp"

namespace vegas {

enum class Model { Ising, Heisenberg };
enum class Algorithm { Metropolis, Wolff };

struct RelaxStage {
    uint64_t steps = 1000;
    double temperature = 3.0;
};
struct CoolDownStage {
    double max_temperature = 3.0, min_temperature = 0.05, cool_rate = 0.1;
    uint64_t relax = 1000, steps = 20000;
};
struct HysteresisStage {
    uint64_t steps = 1000, relax = 1000;
    double temperature = 3.0, max_field = 1.0, field_step = 0.1;
};
using InputStage = std::variant<RelaxStage, CoolDownStage, HysteresisStage>;

struct StateOutput {
    std::string path;
    uint64_t frequency = 1000;
};
struct Output {
    std::optional<std::string> observables;
    std::optional<StateOutput> state;
};

struct Input {
    Model model = Model::Ising;
    Algorithm algorithm = Algorithm::Metropolis;
    std::optional<double> exchange;
    // sample
    bool unitcell_is_path = false;
    std::string unitcell = "sc";  // name, or the path
    uint64_t size[3] = {1, 1, 1};
    bool pbc[3] = {true, true, true};
    std::vector<InputStage> stages;
    std::optional<Output> output;

    // Input::default (input.rs:176-190): Relax and CoolDown with their
    // defaults, both output files.
    static Input defaults() {
        Input in;
        in.stages = {RelaxStage{}, CoolDownStage{}};
        Output o;
        o.observables = "./output.parquet";
        o.state = StateOutput{"./state.parquet", 1000};
        in.output = o;
        return in;
    }

    static Input from_toml(const std::string& text);
    std::string to_toml() const;
    Lattice lattice() const;
    // Runs every stage (input.rs:264-294); Stat lines go to `out`.
    template <class Rng>
    void run(Rng& rng, std::ostream& out) const;

  private:
    template <class S, class I, class Rng>
    void run_with_spin(Rng& rng, std::ostream& out) const;
};

namespace input_detail {

[[noreturn]] inline void fail(const std::string& what) {
    throw Error(ErrorKind::InvalidArgument, "toml deserialization error: " + what);
}
inline const toml::Value* find(const toml::Table& t, const char* key) {
    auto it = t.find(key);
    return it == t.end() ? nullptr : &it->second;
}
inline const toml::Value& need(const toml::Table& t, const char* key, const char* where) {
    const toml::Value* v = find(t, key);
    if (!v) fail(std::string("missing field `") + key + "` in " + where);
    return *v;
}
inline const toml::Table& as_table(const toml::Value& v, const char* what) {
    if (v.type != toml::Value::Type::Table)
        fail(std::string("invalid type: ") + v.type_name() + ", expected a table for " + what);
    return *v.tab;
}
inline std::string as_string(const toml::Value& v, const char* what) {
    if (v.type != toml::Value::Type::String)
        fail(std::string("invalid type: ") + v.type_name() + ", expected a string for " + what);
    return v.s;
}
inline double as_f64(const toml::Value& v, const char* what) {
    if (v.type == toml::Value::Type::Float) return v.f;
    if (v.type == toml::Value::Type::Integer) return (double)v.i;
    fail(std::string("invalid type: ") + v.type_name() + ", expected f64 for " + what);
}
inline uint64_t as_usize(const toml::Value& v, const char* what) {
    if (v.type != toml::Value::Type::Integer)
        fail(std::string("invalid type: ") + v.type_name() + ", expected usize for " + what);
    if (v.i < 0) fail(std::string("invalid value: integer `") + std::to_string(v.i) +
                      "`, expected usize for " + what);
    return (uint64_t)v.i;
}
inline bool as_bool(const toml::Value& v, const char* what) {
    if (v.type != toml::Value::Type::Boolean)
        fail(std::string("invalid type: ") + v.type_name() + ", expected a boolean for " + what);
    return v.b;
}

}  // namespace input_detail

inline Input Input::from_toml(const std::string& text) {
    using namespace input_detail;
    const toml::Table root = toml::parse(text);
    Input in;
    const std::string model = as_string(need(root, "model", "Input"), "model");
    if (model == "Ising")
        in.model = Model::Ising;
    else if (model == "Heisenberg")
        in.model = Model::Heisenberg;
    else
        fail("unknown variant `" + model + "`, expected `Ising` or `Heisenberg`");
    const std::string alg = as_string(need(root, "algorithm", "Input"), "algorithm");
    if (alg == "Metropolis")
        in.algorithm = Algorithm::Metropolis;
    else if (alg == "Wolff")
        in.algorithm = Algorithm::Wolff;
    else
        fail("unknown variant `" + alg + "`, expected `Metropolis` or `Wolff`");
    if (const toml::Value* e = find(root, "exchange")) in.exchange = as_f64(*e, "exchange");

    const toml::Table& sample = as_table(need(root, "sample", "Input"), "sample");
    const toml::Table& uc = as_table(need(sample, "unitcell", "Sample"), "unitcell");
    if (uc.size() != 1) fail("unitcell must be one of {name = ...} or {path = ...}");
    if (const toml::Value* n = find(uc, "name")) {
        in.unitcell = as_string(*n, "unitcell.name");
        if (in.unitcell != "sc" && in.unitcell != "bcc" && in.unitcell != "fcc")
            fail("unknown variant `" + in.unitcell + "`, expected one of `sc`, `bcc`, `fcc`");
    } else if (const toml::Value* p = find(uc, "path")) {
        in.unitcell_is_path = true;
        in.unitcell = as_string(*p, "unitcell.path");
    } else {
        fail("unknown variant `" + uc.begin()->first + "`, expected `name` or `path`");
    }
    const toml::Table& size = as_table(need(sample, "size", "Sample"), "size");
    const toml::Table& pbc = as_table(need(sample, "pbc", "Sample"), "pbc");
    const char* axes[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; ++a) {
        in.size[a] = as_usize(need(size, axes[a], "UnitCellSize"), axes[a]);
        in.pbc[a] = as_bool(need(pbc, axes[a], "PeriodicBoundaryConditions"), axes[a]);
    }

    const toml::Value& st = need(root, "stages", "Input");
    if (st.type != toml::Value::Type::Array) fail("invalid type: expected a sequence for stages");
    for (const toml::Value& sv : *st.arr) {
        const toml::Table& s = as_table(sv, "a stage");
        const std::string prog = as_string(need(s, "program", "Stage"), "program");
        if (prog == "Relax") {
            RelaxStage r;
            r.steps = as_usize(need(s, "steps", "Relax"), "steps");
            r.temperature = as_f64(need(s, "temperature", "Relax"), "temperature");
            in.stages.push_back(r);
        } else if (prog == "CoolDown") {
            CoolDownStage c;
            c.max_temperature = as_f64(need(s, "max_temperature", "CoolDown"), "max_temperature");
            c.min_temperature = as_f64(need(s, "min_temperature", "CoolDown"), "min_temperature");
            c.cool_rate = as_f64(need(s, "cool_rate", "CoolDown"), "cool_rate");
            c.relax = as_usize(need(s, "relax", "CoolDown"), "relax");
            c.steps = as_usize(need(s, "steps", "CoolDown"), "steps");
            in.stages.push_back(c);
        } else if (prog == "Hysteresis") {
            HysteresisStage h;
            h.steps = as_usize(need(s, "steps", "HysteresisLoop"), "steps");
            h.relax = as_usize(need(s, "relax", "HysteresisLoop"), "relax");
            h.temperature = as_f64(need(s, "temperature", "HysteresisLoop"), "temperature");
            h.max_field = as_f64(need(s, "max_field", "HysteresisLoop"), "max_field");
            h.field_step = as_f64(need(s, "field_step", "HysteresisLoop"), "field_step");
            in.stages.push_back(h);
        } else {
            fail("unknown variant `" + prog + "`, expected one of `Relax`, `CoolDown`, `Hysteresis`");
        }
    }

    if (const toml::Value* ov = find(root, "output")) {
        const toml::Table& o = as_table(*ov, "output");
        Output out;
        if (const toml::Value* p = find(o, "observables")) out.observables = as_string(*p, "observables");
        if (const toml::Value* sv = find(o, "state")) {
            const toml::Table& s = as_table(*sv, "output.state");
            StateOutput so;
            so.path = as_string(need(s, "path", "StateOutput"), "path");
            so.frequency = as_usize(need(s, "frequency", "StateOutput"), "frequency");
            out.state = so;
        }
        in.output = out;
    }
    return in;
}

inline std::string Input::to_toml() const {
    using toml::float_literal;
    using toml::quote;
    std::ostringstream o;
    o << "model = " << quote(model == Model::Ising ? "Ising" : "Heisenberg") << "\n";
    o << "algorithm = " << quote(algorithm == Algorithm::Metropolis ? "Metropolis" : "Wolff")
      << "\n";
    if (exchange) o << "exchange = " << float_literal(*exchange) << "\n";
    o << "\n[sample.unitcell]\n" << (unitcell_is_path ? "path" : "name") << " = "
      << quote(unitcell) << "\n";
    o << "\n[sample.size]\nx = " << size[0] << "\ny = " << size[1] << "\nz = " << size[2] << "\n";
    o << "\n[sample.pbc]\nx = " << (pbc[0] ? "true" : "false") << "\ny = "
      << (pbc[1] ? "true" : "false") << "\nz = " << (pbc[2] ? "true" : "false") << "\n";
    for (const InputStage& s : stages) {
        o << "\n[[stages]]\n";
        if (auto r = std::get_if<RelaxStage>(&s)) {
            o << "program = \"Relax\"\nsteps = " << r->steps
              << "\ntemperature = " << float_literal(r->temperature) << "\n";
        } else if (auto c = std::get_if<CoolDownStage>(&s)) {
            o << "program = \"CoolDown\"\nmax_temperature = " << float_literal(c->max_temperature)
              << "\nmin_temperature = " << float_literal(c->min_temperature)
              << "\ncool_rate = " << float_literal(c->cool_rate) << "\nrelax = " << c->relax
              << "\nsteps = " << c->steps << "\n";
        } else if (auto h = std::get_if<HysteresisStage>(&s)) {
            o << "program = \"Hysteresis\"\nsteps = " << h->steps << "\nrelax = " << h->relax
              << "\ntemperature = " << float_literal(h->temperature)
              << "\nmax_field = " << float_literal(h->max_field)
              << "\nfield_step = " << float_literal(h->field_step) << "\n";
        }
    }
    if (output) {
        o << "\n[output]\n";
        if (output->observables) o << "observables = " << quote(*output->observables) << "\n";
        if (output->state)
            o << "\n[output.state]\npath = " << quote(output->state->path)
              << "\nfrequency = " << output->state->frequency << "\n";
    }
    return o.str();
}

// Input::lattice (input.rs:296-322): the named unit cell expanded by size,
// with drop_x/y/z on the open axes.
inline Lattice Input::lattice() const {
    if (unitcell_is_path)
        throw Error(ErrorKind::NotImplemented, "unit cells read from files are not supported");
    Lattice l = unitcell == "sc" ? Lattice::sc() : (unitcell == "bcc" ? Lattice::bcc() : Lattice::fcc());
    l = l.expand((uint32_t)size[0], (uint32_t)size[1], (uint32_t)size[2]);
    if (!pbc[0]) l = l.drop_x();
    if (!pbc[1]) l = l.drop_y();
    if (!pbc[2]) l = l.drop_z();
    return l;
}

// Input::run (input.rs:347-367): Ising + Metropolis runs the flip
// integrator, Ising + Wolff the cluster one, Heisenberg + Metropolis the
// fresh-proposal one; Heisenberg + Wolff is NotImplementedError, as in the
// reference.
template <class Rng>
void Input::run(Rng& rng, std::ostream& out) const {
    if (model == Model::Ising && algorithm == Algorithm::Metropolis)
        run_with_spin<IsingSpin, MetropolisFlipIntegrator>(rng, out);
    else if (model == Model::Ising)
        run_with_spin<IsingSpin, WolffIntegrator>(rng, out);
    else if (algorithm == Algorithm::Metropolis)
        run_with_spin<HeisenbergSpin, MetropolisIntegrator>(rng, out);
    else
        throw Error(ErrorKind::NotImplemented);
}

// Input::run_with_spin (input.rs:264-294): Compound(Exchange, Zeeman) on the
// lattice, a StatSensor on `out` plus the output sensors, thermostat 2.8 with
// no field, a random initial state, then every stage in order.
template <class S, class I, class Rng>
void Input::run_with_spin(Rng& rng, std::ostream& out) const {
    const Lattice l = lattice();
    auto h = hamiltonian(Exchange::from_lattice(exchange.value_or(1.0), l), Zeeman<S>());
    using H = decltype(h);
    typename Machine<H, I, S>::Instruments ins;
    ins.emplace_back(new StatSensor<H, S>(out));
    if (output && output->observables) ins.emplace_back(new ObservableSensor<H, S>(*output->observables));
    if (output && output->state)
        ins.emplace_back(new StateSensor<H, S>(output->state->path, (size_t)output->state->frequency));
    auto state = State<S>::rand_with_size(rng, (size_t)l.num_sites());
    Machine<H, I, S> m(Thermostat<S>(2.8, Field<S>::zero()), std::move(h), I(), std::move(ins),
                       std::move(state));
    for (const InputStage& s : stages) {
        if (auto r = std::get_if<RelaxStage>(&s))
            Relax((size_t)r->steps, r->temperature).run(rng, m);
        else if (auto c = std::get_if<CoolDownStage>(&s))
            CoolDown(c->max_temperature, c->min_temperature, c->cool_rate, (size_t)c->relax,
                     (size_t)c->steps)
                .run(rng, m);
        else if (auto hs = std::get_if<HysteresisStage>(&s))
            HysteresisLoop((size_t)hs->steps, (size_t)hs->relax, hs->temperature, hs->max_field,
                           hs->field_step)
                .run(rng, m);
    }
}

}  // namespace vegas
