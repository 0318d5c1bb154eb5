// Checkpoint and resume of a Machine between the stages of a program.
//
// The reference has none: a run that stops starts over (its StateSensor
// writes spin snapshots that nothing reads back, src/instrument.rs:340-350,
// and the Pcg64 state is never saved).  Here a chain's randomness is
// counter-based -- Philox keyed by the chain's seed, its sweep counter and
// the site -- so a chain is fully described by its seed, the number of
// sweeps it has run, its spins and its thermostat.  A checkpoint taken
// between two stages holds those and the index of the next stage; a
// Machine restored from it and run from that stage (run_stages) repeats the
// uninterrupted run bit for bit: the same spins, per-sweep energies and
// StatSensor lines.  Instruments start every phase afresh, so no
// accumulator state crosses a stage boundary.
//
// File (little-endian): the 8 bytes "VGCKPT01"; u32 spin kind (VG_SPIN_*);
// u32 started (1 once the chain exists: seed and sweeps are meaningful);
// u64 seed; u64 sweeps; u64 next stage; f64 temperature; f64 field
// magnitude; f64 field orientation x, y, z; u64 sites; then the spins, one
// int8 (+1 / -1) per site for Ising or three f64 (x, y, z) per site for
// Heisenberg.  Written to <path>.tmp and renamed into place, so a file at
// <path> is always whole.
#pragma once

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "machine.hpp"
#include "program.hpp"

namespace vegas {

namespace ckpt_detail {

constexpr char kMagic[8] = {'V', 'G', 'C', 'K', 'P', 'T', '0', '1'};

inline void put(std::string& out, const void* p, size_t n) {
    out.append(static_cast<const char*>(p), n);
}

struct Reader {
    const std::string& s;
    size_t at = 0;
    void get(void* p, size_t n) {
        if (at + n > s.size()) throw Error(ErrorKind::InvalidArgument, "checkpoint file truncated");
        std::memcpy(p, s.data() + at, n);
        at += n;
    }
};

inline void spin_vector(const IsingSpin& s, double v[3]) {
    v[0] = 0.0;
    v[1] = 0.0;
    v[2] = s.sz();
}
inline void spin_vector(const HeisenbergSpin& s, double v[3]) {
    v[0] = s.sx();
    v[1] = s.sy();
    v[2] = s.sz();
}
inline IsingSpin spin_of(IsingSpin*, const double v[3]) {
    return v[2] < 0.0 ? IsingSpin::down() : IsingSpin::up();
}
inline HeisenbergSpin spin_of(HeisenbergSpin*, const double v[3]) {
    return HeisenbergSpin(v[0], v[1], v[2]);
}

inline void put_spins(std::string& out, const State<IsingSpin>& st) {
    for (size_t i = 0; i < st.len(); ++i) {
        const int8_t v = (int8_t)st.at(i).value();
        put(out, &v, 1);
    }
}
inline void put_spins(std::string& out, const State<HeisenbergSpin>& st) {
    for (size_t i = 0; i < st.len(); ++i) {
        const double v[3] = {st.at(i).sx(), st.at(i).sy(), st.at(i).sz()};
        put(out, v, sizeof v);
    }
}
inline State<IsingSpin> get_spins(Reader& r, size_t n, IsingSpin*) {
    std::vector<IsingSpin> v;
    v.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        int8_t b;
        r.get(&b, 1);
        if (b != 1 && b != -1) throw Error(ErrorKind::InvalidArgument, "checkpoint: bad Ising spin");
        v.push_back(IsingSpin(b > 0 ? IsingSpin::Up : IsingSpin::Down));
    }
    return State<IsingSpin>(std::move(v));
}
inline State<HeisenbergSpin> get_spins(Reader& r, size_t n, HeisenbergSpin*) {
    std::vector<HeisenbergSpin> v;
    v.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        double x[3];
        r.get(x, sizeof x);
        v.emplace_back(x[0], x[1], x[2]);
    }
    return State<HeisenbergSpin>(std::move(v));
}

}  // namespace ckpt_detail

// The bytes of a checkpoint file.
template <class S>
std::string encode_checkpoint(const Checkpoint<S>& ck) {
    using namespace ckpt_detail;
    std::string out;
    put(out, kMagic, 8);
    const uint32_t kind = (uint32_t)S::kind, started = ck.started ? 1u : 0u;
    put(out, &kind, 4);
    put(out, &started, 4);
    put(out, &ck.seed, 8);
    put(out, &ck.sweeps, 8);
    put(out, &ck.next_stage, 8);
    const double t = ck.thermostat.temperature(), h = ck.thermostat.field().magnitude();
    double dir[3];
    spin_vector(ck.thermostat.field().orientation(), dir);
    put(out, &t, 8);
    put(out, &h, 8);
    put(out, dir, sizeof dir);
    const uint64_t n = ck.state.len();
    put(out, &n, 8);
    put_spins(out, ck.state);
    return out;
}

template <class S>
Checkpoint<S> decode_checkpoint(const std::string& bytes) {
    using namespace ckpt_detail;
    Reader r{bytes};
    char magic[8];
    r.get(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0)
        throw Error(ErrorKind::InvalidArgument, "not a checkpoint file");
    uint32_t kind, started;
    r.get(&kind, 4);
    r.get(&started, 4);
    if (kind != (uint32_t)S::kind)
        throw Error(ErrorKind::InvalidArgument, "checkpoint holds another spin kind");
    if (started > 1) throw Error(ErrorKind::InvalidArgument, "checkpoint: bad header");
    Checkpoint<S> ck;
    ck.started = started == 1;
    r.get(&ck.seed, 8);
    r.get(&ck.sweeps, 8);
    r.get(&ck.next_stage, 8);
    double t, h, dir[3];
    r.get(&t, 8);
    r.get(&h, 8);
    r.get(dir, sizeof dir);
    ck.thermostat = Thermostat<S>(t, Field<S>(spin_of((S*)nullptr, dir), h));
    uint64_t n;
    r.get(&n, 8);
    const size_t per = S::kind == VG_SPIN_ISING ? 1 : 24;
    if (n > (bytes.size() - r.at) / per)
        throw Error(ErrorKind::InvalidArgument, "checkpoint file truncated");
    ck.state = get_spins(r, (size_t)n, (S*)nullptr);
    if (r.at != bytes.size()) throw Error(ErrorKind::InvalidArgument, "checkpoint: trailing bytes");
    return ck;
}

// Writes the checkpoint to <path>.tmp, then renames it to path.
template <class S>
void save_checkpoint(const std::string& path, const Checkpoint<S>& ck) {
    const std::string bytes = encode_checkpoint(ck);
    const std::string tmp = path + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw Error(ErrorKind::Io, "cannot write " + tmp);
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    if (std::fclose(f) != 0 || !ok) {
        std::remove(tmp.c_str());
        throw Error(ErrorKind::Io, "cannot write " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw Error(ErrorKind::Io, "cannot rename " + tmp + " to " + path);
    }
}

template <class S>
Checkpoint<S> load_checkpoint(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(ErrorKind::Io, "cannot read " + path);
    std::string bytes;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.append(buf, n);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) throw Error(ErrorKind::Io, "cannot read " + path);
    return decode_checkpoint<S>(bytes);
}

// Runs stages[first..] of a program's schedule on m as the program's run()
// does -- the thermostat's temperature (and field, for stages that set it),
// then relax_for and measure_for -- and calls after_stage(i + 1) once
// stage i is done, where a caller may take m.checkpoint(i + 1).
template <class Rng, class H, class I, class S, class F>
void run_stages(Rng& rng, Machine<H, I, S>& m, const std::vector<Stage>& stages, size_t first,
                F&& after_stage) {
    if (first > stages.size())
        throw Error(ErrorKind::InvalidArgument, "resume stage past the program's end");
    for (size_t i = first; i < stages.size(); ++i) {
        const Stage& st = stages[i];
        Thermostat<S> th = m.thermostat().with_temperature(st.temperature);
        if (st.sets_field) th = th.with_field(Field<S>(S::up(), st.field));
        m.set_thermostat(th);
        m.relax_for(rng, st.relax);
        if (!st.relax_only) m.measure_for(rng, st.measure);
        after_stage((uint64_t)(i + 1));
    }
}

}  // namespace vegas
