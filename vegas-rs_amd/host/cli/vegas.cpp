// The `vegas` command line (reference src/main.rs): `vegas input` prints the
// default input, `vegas bench <ising|heisenberg> <length> [--seed N]` runs
// the reference's benchmark program, `vegas run <input.toml|-> [--seed N]`
// runs a simulation from a TOML input.  Sweeps run on the GPU engine; the
// host seeds come from a 64-bit Mersenne Twister seeded with --seed (random
// without), so a seed gives the same run every time, but not the
// reference's PCG64 trajectory.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <random>
#include <sstream>
#include <string>

#include "vegas/input.hpp"
#include "vegas/vegas.hpp"

using namespace vegas;

namespace {

const char* kUsage =
    "Usage: vegas <COMMAND>\n\n"
    "Commands:\n"
    "  input  Print the default input\n"
    "  bench  Run benchmarks\n"
    "  run    Run a simulation from an input file\n"
    "  help   Print this message\n\n"
    "Options:\n"
    "  -h, --help     Print help\n"
    "  -V, --version  Print version\n";

[[noreturn]] void usage_error(const std::string& msg) {
    std::cerr << "error: " << msg << "\n\n" << kUsage;
    std::exit(2);
}

uint64_t parse_u64(const std::string& s, const char* what) {
    if (s.empty() || s.find_first_not_of("0123456789") != std::string::npos)
        usage_error(std::string("invalid value '") + s + "' for " + what);
    errno = 0;
    const unsigned long long v = std::strtoull(s.c_str(), nullptr, 10);
    if (errno) usage_error(std::string("invalid value '") + s + "' for " + what);
    return v;
}

// Positional arguments and an optional -s/--seed of a subcommand.
void split_args(int argc, char** argv, int from, std::vector<std::string>& pos,
                std::optional<uint64_t>& seed) {
    for (int k = from; k < argc; ++k) {
        const std::string a = argv[k];
        if (a == "-s" || a == "--seed") {
            if (k + 1 >= argc) usage_error("a value is required for '--seed <SEED>'");
            seed = parse_u64(argv[++k], "'--seed <SEED>'");
        } else if (a.rfind("--seed=", 0) == 0) {
            seed = parse_u64(a.substr(7), "'--seed <SEED>'");
        } else if (a.size() > 1 && a[0] == '-' && a != "-") {
            usage_error("unexpected argument '" + a + "'");
        } else {
            pos.push_back(a);
        }
    }
}

std::mt19937_64 make_rng(const std::optional<uint64_t>& seed) {
    if (seed) return std::mt19937_64(*seed);
    std::random_device rd;
    return std::mt19937_64(((uint64_t)rd() << 32) ^ rd());
}

// bench_model (main.rs:24-60): sc length^3, Exchange J = 1, fresh-proposal
// Metropolis, CoolDown from 5.0 (Ising) or 2.5 (Heisenberg) in steps of
// 0.05 with the default relax/steps, Stat lines on stdout.
template <class S>
void bench(uint64_t length, std::mt19937_64& rng, double max_temperature) {
    const Lattice l = Lattice::sc().expand_all((uint32_t)length);
    using H = Exchange;
    typename Machine<H, MetropolisIntegrator, S>::Instruments ins;
    ins.emplace_back(new StatSensor<H, S>(std::cout));
    auto state = State<S>::rand_with_size(rng, (size_t)l.num_sites());
    Machine<H, MetropolisIntegrator, S> m(Thermostat<S>(2.8, Field<S>::zero()),
                                          Exchange::from_lattice(1.0, l), MetropolisIntegrator(),
                                          std::move(ins), std::move(state));
    CoolDown().set_max_temperature(max_temperature).set_cool_rate(0.05).run(rng, m);
}

int run_main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << kUsage;
        return 2;
    }
    const std::string cmd = argv[1];
    if (cmd == "-h" || cmd == "--help" || cmd == "help") {
        std::cout << kUsage;
        return 0;
    }
    if (cmd == "-V" || cmd == "--version") {
        std::cout << "vegas " << vg_version() << "\n";
        return 0;
    }
    std::vector<std::string> pos;
    std::optional<uint64_t> seed;
    split_args(argc, argv, 2, pos, seed);
    if (cmd == "input") {
        if (!pos.empty()) usage_error("unexpected argument '" + pos[0] + "'");
        std::cout << Input::defaults().to_toml() << "\n";
        return 0;
    }
    if (cmd == "bench") {
        if (pos.size() != 2) usage_error("bench needs <MODEL> <LENGTH>");
        const uint64_t length = parse_u64(pos[1], "'<LENGTH>'");
        std::mt19937_64 rng = make_rng(seed);
        if (pos[0] == "ising")
            bench<IsingSpin>(length, rng, 5.0);
        else if (pos[0] == "heisenberg")
            bench<HeisenbergSpin>(length, rng, 2.5);
        else
            usage_error("invalid value '" + pos[0] +
                        "' for '<MODEL>' [possible values: ising, heisenberg]");
        return 0;
    }
    if (cmd == "run") {
        if (pos.size() != 1) usage_error("run needs <INPUT>");
        std::string text;
        if (pos[0] == "-") {
            text.assign(std::istreambuf_iterator<char>(std::cin), std::istreambuf_iterator<char>());
        } else {
            std::ifstream f(pos[0], std::ios::binary);
            if (!f) throw Error(ErrorKind::Io, "std io error: cannot open " + pos[0]);
            text.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        }
        const Input in = Input::from_toml(text);
        std::mt19937_64 rng = make_rng(seed);
        in.run(rng, std::cout);
        return 0;
    }
    usage_error("unrecognized subcommand '" + cmd + "'");
}

}  // namespace

// VegasError's display (error.rs:12-27): program errors and I/O errors are
// wrapped ("program error: ...", "io error: ..."); TOML errors carry their
// own prefix; NotImplementedError is "not implemented error".
std::string describe(const Error& e) {
    switch (e.kind()) {
        case ErrorKind::TemperatureMaxLessThanMin:
        case ErrorKind::NoSteps:
        case ErrorKind::ZeroTemperature:
        case ErrorKind::ZeroCoolRate:
        case ErrorKind::ZeroField:
        case ErrorKind::ZeroFieldStep: return std::string("program error: ") + e.what();
        case ErrorKind::Io: return std::string("io error: ") + e.what();
        default: return e.what();
    }
}

int main(int argc, char** argv) {
    try {
        return run_main(argc, argv);
    } catch (const Error& e) {
        std::cout.flush();
        std::cerr << "Error: " << describe(e) << "\n";
        return 1;
    } catch (const std::exception& e) {
        std::cout.flush();
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
}
