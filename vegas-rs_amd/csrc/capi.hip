// C ABI of the engine (declared in include/vegas_hip.h): contexts, batches of
// chains packed in per-kernel arenas, the sweep-block driver (plain launches
// or a replayed hipGraph), and the host side of the observables.
#include "engine.hpp"

using vg::ChainDev;

namespace vgi {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int coordination(int kind) { return kind == VG_LATTICE_SC ? 6 : (kind == VG_LATTICE_BCC ? 8 : 12); }

int validate_desc(const vg_chain_desc& d) {
    const vg_lattice& L = d.lattice;
    if (L.kind == VG_LATTICE_GRAPH) {
        // Exchange::new(matrix): int8 Ising or Heisenberg, Metropolis moves
        if (d.spin != VG_SPIN_ISING && d.spin != VG_SPIN_HEISENBERG)
            return fail(VG_ERR_NOT_IMPLEMENTED, "graph chains take int8 Ising or Heisenberg spins");
        if (d.proposal != VG_PROPOSAL_FLIP && d.proposal != VG_PROPOSAL_RANDOM)
            return fail(VG_ERR_NOT_IMPLEMENTED, "graph chains take flip or random proposals");
        if (!(d.hamiltonian.terms & VG_TERM_EXCHANGE))
            return fail(VG_ERR_INVALID_ARGUMENT, "a graph chain's Hamiltonian needs its Exchange term");
    } else if (L.kind != VG_LATTICE_SC && L.kind != VG_LATTICE_BCC && L.kind != VG_LATTICE_FCC) {
        return fail(VG_ERR_UNSUPPORTED_LATTICE, "unknown lattice kind %d", L.kind);
    }
    if (L.kind != VG_LATTICE_GRAPH) {
        if (L.lx == 0 || L.ly == 0 || L.lz == 0)
            return fail(VG_ERR_INVALID_ARGUMENT, "lattice dimensions must be positive");
        if (L.pbc & ~7u) return fail(VG_ERR_INVALID_ARGUMENT, "unknown pbc bits 0x%x", L.pbc);
        if (L.kind == VG_LATTICE_SC) {
            if (L.lx % 2)
                return fail(VG_ERR_UNSUPPORTED_LATTICE,
                            "sc lattice needs an even x length for the two-colour layout (lx=%u)",
                            L.lx);
            const uint32_t dims[3] = {L.lx, L.ly, L.lz};
            for (int a = 0; a < 3; ++a)
                if ((L.pbc >> a) & 1u && dims[a] % 2)
                    return fail(VG_ERR_UNSUPPORTED_LATTICE,
                                "periodic sc axis %c has odd length %u: not bipartite", "xyz"[a],
                                dims[a]);
        }  // bcc, fcc: the sublattices are the colours for any sizes and boundaries
        const uint64_t cells = (uint64_t)L.lx * L.ly * L.lz;
        if ((uint64_t)L.ly * L.lz > 0xFFFFFFFFull || cells > (1ull << 34))
            return fail(VG_ERR_INVALID_ARGUMENT, "lattice too large");
    }
    if (d.spin != VG_SPIN_ISING && d.spin != VG_SPIN_HEISENBERG && d.spin != VG_SPIN_ISING_BITS)
        return fail(VG_ERR_INVALID_ARGUMENT, "unknown spin kind %d", d.spin);
    if (d.proposal != VG_PROPOSAL_FLIP && d.proposal != VG_PROPOSAL_RANDOM &&
        d.proposal != VG_PROPOSAL_WOLFF)
        return fail(VG_ERR_INVALID_ARGUMENT, "unknown proposal %d", d.proposal);
    if (d.proposal == VG_PROPOSAL_WOLFF) {
        // integrator.rs:172-230 implements Wolff for Ising spins only
        if (d.spin != VG_SPIN_ISING)
            return fail(VG_ERR_NOT_IMPLEMENTED, "Wolff cluster steps run on int8 Ising chains");
        // site ids and cell coordinates by 31-bit magic division
        if (sites_of(L) >= (1ull << 31))
            return fail(VG_ERR_INVALID_ARGUMENT, "Wolff chains need fewer than 2^31 sites");
    }
    const vg_hamiltonian& H = d.hamiltonian;
    if (H.terms & ~15u) return fail(VG_ERR_INVALID_ARGUMENT, "unknown term bits 0x%x", H.terms);
    if (H.terms == 0) return fail(VG_ERR_INVALID_ARGUMENT, "hamiltonian has no terms");
    const int nterms = __builtin_popcount(H.terms);
    if (nterms > 1 && !H.compound)
        return fail(VG_ERR_INVALID_ARGUMENT, "several terms need compound = 1");
    if (nterms == 1 && H.compound)
        return fail(VG_ERR_INVALID_ARGUMENT, "a Compound needs at least two terms");
    return VG_OK;
}

// Acceptance thresholds of an Ising chain for its current thermostat.
// Class (b, S): current spin s = 2b-1, neighbour sum sigma = S - z.
// dE = ((2*J)*s)*sigma - ((2*|h|)*hz)*s, evaluated in f64 left to right;
// dE <= 0 is always accepted (the reference accepts dE < 0 outright and
// dE = 0 with probability 1), otherwise the draw k (uniform 32-bit) is
// accepted iff k < ceil(exp(-dE/T) * 2^32).
void ising_thresholds(const ChainHost& ch, ChainDev& d) {
    const vg_hamiltonian& H = ch.desc.hamiltonian;
    const int z = coordination(ch.desc.lattice.kind);
    const double J = (H.terms & VG_TERM_EXCHANGE) ? H.exchange : 0.0;
    const double hm = (H.terms & VG_TERM_ZEEMAN) ? std::fabs(ch.field_mag) : 0.0;
    const double hz = ch.field_dir[2] >= 0.0 ? 1.0 : -1.0;
    const double T = ch.temperature;
    for (int b = 0; b < 2; ++b) {
        d.always[b] = 0;
        const double s = b ? 1.0 : -1.0;
        for (int S = 0; S < (int)vg::kThrS; ++S) {
            d.thr[b][S] = 0;
            if (S > 2 * z) continue;
            const double sigma = (double)(S - z);
            const double dE = 2.0 * J * s * sigma - 2.0 * hm * hz * s;
            if (dE <= 0.0) {
                d.always[b] |= 1u << S;
                continue;
            }
            const double th = std::ceil(std::ldexp(std::exp(-dE / T), 32));
            if (th >= 4294967296.0)
                d.always[b] |= 1u << S;
            else
                d.thr[b][S] = (uint32_t)th;
        }
    }
    // Class plans of the multispin sc pass (z = 6): for a site with g absent
    // neighbour slots, p = 6 - g present and a of them aligned with its bit b,
    // the up-neighbour count is c = a (b = 1) or p - a (b = 0) and the class
    // S = 2c + g.  Classes always accepted are left out.
    std::memset(d.plan, 0, sizeof d.plan);
    if (z == 6)
        for (uint32_t g = 0; g < vg::kPlanG; ++g)
            for (uint32_t a = 0; a <= 6 - g; ++a) {
                uint32_t slot[2] = {0, 0};
                for (uint32_t b = 0; b < 2; ++b) {
                    const uint32_t S = 2 * (b ? a : 6 - g - a) + g;
                    if ((d.always[b] >> S) & 1u) continue;
                    const uint32_t th = d.thr[b][S];
                    slot[b] = (1u << (3 + b)) | ((th != 0 ? 1u : 0u) << 5) | ((th >> 24) << 8);
                }
                if (slot[0] && slot[1] && (slot[0] >> 5) == (slot[1] >> 5)) {
                    d.plan[g][a][0] = slot[0] | slot[1];
                } else {
                    d.plan[g][a][0] = slot[0] ? slot[0] : slot[1];
                    d.plan[g][a][1] = slot[0] ? slot[1] : 0;
                }
            }
}

// m, s with n / d == (n * m) >> s for every n < 2^31: s = 31 + ceil(log2 d),
// m = ceil(2^s / d) < 2^32; the error n*(m - 2^s/d)/2^s < 1/d never carries.
void magic_div(uint32_t d, uint32_t& m, uint32_t& s) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    s = 31 + l;
    m = (uint32_t)(((1ull << s) + d - 1) / d);
}

void heis_params(const ChainHost& ch, ChainDev& d) {
    const vg_hamiltonian& H = ch.desc.hamiltonian;
    d.jn = (H.terms & VG_TERM_EXCHANGE) ? (float)(-H.exchange) : 0.0f;
    d.kk = (H.terms & VG_TERM_ANISOTROPY) ? (float)H.anisotropy : 0.0f;
    d.hm = (H.terms & VG_TERM_ZEEMAN) ? (float)std::fabs(ch.field_mag) : 0.0f;
    d.beta = (float)(1.0 / ch.temperature);
    d.ax = (float)H.anisotropy_axis[0];
    d.ay = (float)H.anisotropy_axis[1];
    d.az = (float)H.anisotropy_axis[2];
    d.fx = (float)ch.field_dir[0];
    d.fy = (float)ch.field_dir[1];
    d.fz = (float)ch.field_dir[2];
}

// Ising parameters of a graph chain, f64: 1/T, |h| (0 without Zeeman) and the
// sign of the field's orientation (Field of an Ising system is Up or Down).
void graph_ising_params(const ChainHost& ch, ChainDev& d) {
    const vg_hamiltonian& H = ch.desc.hamiltonian;
    d.gbeta = 1.0 / ch.temperature;
    d.ghm = (H.terms & VG_TERM_ZEEMAN) ? std::fabs(ch.field_mag) : 0.0;
    d.ghz = ch.field_dir[2] >= 0.0 ? 1.0 : -1.0;
}

// Wolff bond threshold: a bond is active iff its 32-bit draw is below
// ceil(p 2^32), p = 1 - exp(-2J/T) in f64 (integrator.rs:188-192).
void wolff_threshold(const ChainHost& ch, ChainDev& d) {
    const vg_hamiltonian& H = ch.desc.hamiltonian;
    const double J = (H.terms & VG_TERM_EXCHANGE) ? H.exchange : 0.0;
    const double p = 1.0 - std::exp(-2.0 * J / ch.temperature);
    d.wolff_thr = 0;
    d.wolff_all = 0;
    if (!(p > 0.0)) return;
    const double th = std::ceil(std::ldexp(p, 32));
    if (th >= 4294967296.0)
        d.wolff_all = 1;
    else
        d.wolff_thr = (uint32_t)th;
}

void refresh_params(vg_batch* B, uint32_t i) {
    ChainHost& ch = B->chains[i];
    ChainDev& d = B->groups[ch.group].host[ch.index];
    if (ch.group == kIsingWolff) {
        wolff_threshold(ch, d);
    } else if (ch.group == kIsingGraph) {
        graph_ising_params(ch, d);
    } else if (ch.group == kHeisGraph) {
        heis_params(ch, d);
        d.jn = -1.0f;  // the couplings are inside the field h
    } else if (is_ising(ch.group)) {
        ising_thresholds(ch, d);
    } else {
        heis_params(ch, d);
    }
}

// Ising passes write {mag, bond} per wave (4 waves per workgroup); Heisenberg
// passes write {bond, mx, my, mz, axis2} per workgroup.
int init_chain_dev(const ChainHost& ch, int g, uint32_t slot, ChainDev& d) {
    std::memset(&d, 0, sizeof d);
    const vg_lattice& L = ch.desc.lattice;
    d.seed = ch.desc.seed;
    d.lx = L.lx;
    d.ly = L.ly;
    d.lz = L.lz;
    d.kind = (uint32_t)L.kind;
    d.pbc = L.pbc;
    d.z = (uint32_t)coordination(L.kind);
    d.ncol = colours_of(g);
    d.rows = L.ly * L.lz;
    d.w = L.kind == VG_LATTICE_SC ? L.lx / 2 : L.lx;
    d.proposal = (uint32_t)ch.desc.proposal;
    d.slot = slot;
    d.limit = 0xFFFFFFFFu;
    d.nsites = sites_of(L);
    if (g == kIsingWolff) {  // natural order; no planes or work items
        d.plane = d.nsites;
        magic_div(d.lx, d.lx_m, d.lx_s);
        magic_div(d.ly, d.ly_m, d.ly_s);
        wolff_threshold(ch, d);
        return VG_OK;
    }
    if (is_graph(g)) {  // sections and colour classes are set by the caller
        d.nsites = ch.n_sites;
        d.ncol = 0;
        d.z = 0;
        if (g == kIsingGraph) {
            graph_ising_params(ch, d);
        } else {
            heis_params(ch, d);
            d.jn = -1.0f;
        }
        return VG_OK;
    }
    const uint32_t seg = (g == kIsingSc || g == kIsingBits) ? vg::kScSeg
                         : is_ising(g)                        ? vg::kIsingSeg
                                                              : vg::kHeisSeg;
    const uint32_t align = is_ising(g) ? 64u : 16u;  // 64-byte aligned rows
    if (is_bits(g))  // bytes: whole segments (2 or 4 bytes), 16-byte aligned rows
        d.pitch = (uint32_t)round_up((d.w + seg - 1) / seg * (seg / 8), 16);
    else
        d.pitch = (uint32_t)round_up(d.w, align);
    d.segs = (d.w + seg - 1) / seg;
    d.hy = (d.ly + 1) / 2;
    if ((uint64_t)d.rows * d.segs > 0x7FFFFFFFull - vg::kBlock || (uint64_t)d.rows > 0x7FFFFFFFull)
        return fail(VG_ERR_INVALID_ARGUMENT, "chain %u too large", slot);
    magic_div(d.segs, d.segs_m, d.segs_s);
    magic_div(d.ly, d.ly_m, d.ly_s);
    magic_div(d.hy, d.hy_m, d.hy_s);
    d.plane = (uint64_t)d.rows * d.pitch;
    if (is_ising(g) && d.plane > 0xFFFFF000ull)
        return fail(VG_ERR_UNSUPPORTED_LATTICE,
                    "chain %u: an Ising colour plane of %llu bytes exceeds 4 GiB", slot,
                    (unsigned long long)d.plane);
    if ((g == kHeisSc || g == kHeisBcc || g == kHeisFcc) && 4 * d.plane > 0xFFFFF000ull)
        return fail(VG_ERR_UNSUPPORTED_LATTICE,
                    "chain %u: a Heisenberg colour plane component of %llu bytes exceeds 4 GiB",
                    slot, (unsigned long long)(4 * d.plane));
    if (is_ising(g))
        ising_thresholds(ch, d);
    else
        heis_params(ch, d);
    return VG_OK;
}

uint64_t colour_elems(int g, const ChainDev& d) { return (is_ising(g) ? 1 : 3) * d.plane; }

// Ising lattice passes write {mag, bond} per wave (4 waves); Heisenberg and
// graph passes write {bond, mx, my, mz, axis2} per workgroup.
uint32_t quantities(int group) { return is_ising(group) && !is_graph(group) ? 8u : 5u; }
uint32_t quantity_bytes(int group) { return is_ising(group) && !is_graph(group) ? 4u : 8u; }

// After the int8 planes of an sc chain with shadows were written: rebuild
// its shadows (one thread per shadow word).
void launch_shadow_build(const Group& G, const ChainDev& d, const ChainDev* dp, hipStream_t st) {
    if (!d.shadow) return;
    const uint64_t n = 2ull * d.rows * (d.spitch / 4);
    hipLaunchKernelGGL(vg::ising_shadow_build, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                       dp, (uint8_t*)G.arena);
}

void launch_ising_init(int g, const Group& G, const ChainDev& d, const ChainDev* dp,
                       hipStream_t st) {
    const uint64_t n = (uint64_t)d.ncol * d.plane;  // one thread per plane byte
    const dim3 grid((uint32_t)((n + 255) / 256));
    if (is_bits(g))
        hipLaunchKernelGGL(vg::ising_init_bits, grid, dim3(256), 0, st, dp, (uint8_t*)G.arena);
    else
        hipLaunchKernelGGL(vg::ising_init, grid, dim3(256), 0, st, dp, (uint8_t*)G.arena);
    if (g == kIsingSc) launch_shadow_build(G, d, dp, st);
}

int upload_descriptors(vg_batch* B, int g, int slot) {
    Group& G = B->groups[g];
    if (G.members.empty()) return VG_OK;
    ChainDev* staging = slot ? G.h_staging2 : G.h_staging;
    std::memcpy(staging, G.host.data(), G.host.size() * sizeof(ChainDev));
    VG_HIP(hipMemcpyAsync(G.d_chains, staging, G.host.size() * sizeof(ChainDev),
                          hipMemcpyHostToDevice, B->ctx->stream));
    return VG_OK;
}

constexpr uint32_t kDefaultTile = 4;

template <uint32_t K>
const void* zmarch_fn(int g) {
    switch (g) {
        case kIsingBits: return (const void*)&vg::ising_sc_pass<K, vg::kScBits>;
        case kIsingBcc: return (const void*)&vg::ising_bcc_pass<K, false>;
        case kIsingBccBits: return (const void*)&vg::ising_bcc_pass<K, true>;
        case kIsingFcc: return (const void*)&vg::ising_fcc_pass<0, K, false>;
        case kIsingFccBits: return (const void*)&vg::ising_fcc_pass<0, K, true>;
        default: return (const void*)&vg::ising_sc_pass<K, vg::kScShadow>;
    }
}

// Whether `grid` workgroups of group g's z-marching Ising pass with this tile
// are all resident at once on the current device.
bool all_resident(int g, uint32_t tile, uint32_t grid) {
    const void* fn = tile == 1   ? zmarch_fn<1>(g)
                     : tile == 2 ? zmarch_fn<2>(g)
                     : tile == 8 ? zmarch_fn<8>(g)
                                 : zmarch_fn<4>(g);
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, vg::kBlock, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return false;
    return (uint64_t)grid <= (uint64_t)nb * (uint64_t)cus;
}

// Work decomposition of a group for a tile (Ising sc: z planes per work item):
// items and blocks per chain, the block -> chain map, the grid.  Replays of
// captured graphs bake the grid in, so they are dropped.
int layout_work(vg_batch* B, int g, uint32_t tile) {
    Group& G = B->groups[g];
    std::vector<uint32_t> block_chain;
    for (size_t k = 0; k < G.members.size(); ++k) {
        ChainDev& d = G.host[k];
        uint64_t items;
        if (g == kIsingWolff) {
            items = 1;  // one workgroup per chain
        } else if (is_graph(g)) {
            items = d.items;  // one thread per site of the largest colour class (set at creation)
        } else if (is_zmarch(g)) {
            d.nzc = (d.lz + tile - 1) / tile;
            // sc: rows in natural order; bcc, fcc: the rows of one y parity,
            // then the other (2 * hy rows, the last one past an odd ly idle)
            const bool sc = g == kIsingSc || g == kIsingBits;
            items = (uint64_t)d.nzc * (sc ? d.ly : 2 * d.hy) * d.segs;
        } else {
            items = (uint64_t)d.rows * d.segs;
        }
        if (items > 0x7FFFFFFFull - vg::kBlock)
            return fail(VG_ERR_INVALID_ARGUMENT, "chain %u too large", G.members[k]);
        d.items = (uint32_t)items;
        d.nblocks = (uint32_t)((items + vg::kBlock - 1) / vg::kBlock);
        d.block0 = (uint32_t)block_chain.size();
        for (uint32_t bidx = 0; bidx < d.nblocks; ++bidx) block_chain.push_back((uint32_t)k);
    }
    if (block_chain.size() > G.block_chain_cap) {
        if (G.d_block_chain) quiet(hipFree(G.d_block_chain));
        G.d_block_chain = nullptr;
        G.block_chain_cap = 0;
        VG_HIP(hipMalloc(&G.d_block_chain, block_chain.size() * sizeof(uint32_t)));
        G.block_chain_cap = (uint32_t)block_chain.size();
    }
    VG_HIP(hipMemcpy(G.d_block_chain, block_chain.data(), block_chain.size() * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
    G.tile = tile;
    G.total_blocks = (uint32_t)block_chain.size();
    G.grid = (uint32_t)round_up(G.total_blocks, vg::kNumXcd);
    G.uniform_nb = G.host.empty() ? 0u : G.host[0].nblocks;
    for (const ChainDev& d : G.host)
        if (d.nblocks != G.uniform_nb) G.uniform_nb = 0;
    if (is_zmarch(g)) G.prio = all_resident(g, tile, G.grid) ? 1u : 0u;
    G.partial_bytes_per_sweep = std::max<uint64_t>(
        256, (uint64_t)passes_of(G, g) * G.total_blocks * quantities(g) * quantity_bytes(g));
    for (auto& kv : B->graphs) quiet(hipGraphExecDestroy(kv.second));
    B->graphs.clear();
    return VG_OK;
}

int free_batch(vg_batch* B) {
    if (!B) return VG_OK;
    for (auto& kv : B->graphs) quiet(hipGraphExecDestroy(kv.second));
    B->graphs.clear();
    for (Group& G : B->groups) {
        if (G.d_chains) quiet(hipFree(G.d_chains));
        if (G.h_staging) quiet(hipHostFree(G.h_staging));
        if (G.h_staging2) quiet(hipHostFree(G.h_staging2));
        if (G.d_block_chain) quiet(hipFree(G.d_block_chain));
        if (G.arena && !G.shared_arena) quiet(hipFree(G.arena));
        if (G.partials) quiet(hipFree(G.partials));
    }
    if (B->d_records) quiet(hipFree(B->d_records));
    if (B->h_records) quiet(hipHostFree(B->h_records));
    if (B->h_records2) quiet(hipHostFree(B->h_records2));
    for (hipEvent_t e : {B->ev0, B->ev1, B->ev2, B->ev3, B->done[0], B->done[1]})
        if (e) quiet(hipEventDestroy(e));
    delete B;
    return VG_OK;
}

template <uint32_t MODE>
void launch_sc_mode(const Group& G, hipStream_t st, uint32_t c, uint32_t s) {
    const dim3 grid(G.grid), block(vg::kBlock);
    uint8_t* arena = (uint8_t*)G.arena;
    int32_t* part = (int32_t*)G.partials;
    const uint32_t opts = G.prio | (G.tie_cap << 8);
    switch (G.tile) {
        case 1:
            hipLaunchKernelGGL((vg::ising_sc_pass<1, MODE>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid, opts,
                               G.uniform_nb);
            break;
        case 2:
            hipLaunchKernelGGL((vg::ising_sc_pass<2, MODE>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid, opts,
                               G.uniform_nb);
            break;
        case 8:
            hipLaunchKernelGGL((vg::ising_sc_pass<8, MODE>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid, opts,
                               G.uniform_nb);
            break;
        default:
            hipLaunchKernelGGL((vg::ising_sc_pass<4, MODE>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid, opts,
                               G.uniform_nb);
            break;
    }
}

// One colour pass of an sc Ising group: bit-packed planes, or int8 planes
// with or without shadows.
template <bool BITS>
void launch_ising_pass(const Group& G, hipStream_t st, uint32_t c, uint32_t s) {
    if (BITS)
        launch_sc_mode<vg::kScBits>(G, st, c, s);
    else if (G.shadow)
        launch_sc_mode<vg::kScShadow>(G, st, c, s);
    else
        launch_sc_mode<vg::kScInt8>(G, st, c, s);
}

template <bool BITS>
void launch_bcc_pass(const Group& G, hipStream_t st, uint32_t c, uint32_t s) {
    const dim3 grid(G.grid), block(vg::kBlock);
    uint8_t* arena = (uint8_t*)G.arena;
    int32_t* part = (int32_t*)G.partials;
    switch (G.tile) {
        case 1:
            hipLaunchKernelGGL((vg::ising_bcc_pass<1, BITS>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid,
                               G.prio | (G.tie_cap << 8));
            break;
        case 2:
            hipLaunchKernelGGL((vg::ising_bcc_pass<2, BITS>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid,
                               G.prio | (G.tie_cap << 8));
            break;
        case 8:
            hipLaunchKernelGGL((vg::ising_bcc_pass<8, BITS>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid,
                               G.prio | (G.tie_cap << 8));
            break;
        default:
            hipLaunchKernelGGL((vg::ising_bcc_pass<4, BITS>), grid, block, 0, st, G.d_chains,
                               G.d_block_chain, arena, part, c, s, G.total_blocks, G.grid,
                               G.prio | (G.tie_cap << 8));
            break;
    }
}

template <uint32_t A, uint32_t K, bool BITS>
void launch_ising_fcc(const Group& G, hipStream_t st, uint32_t s) {
    hipLaunchKernelGGL((vg::ising_fcc_pass<A, K, BITS>), dim3(G.grid), dim3(vg::kBlock), 0, st,
                       G.d_chains, G.d_block_chain, (uint8_t*)G.arena, (int32_t*)G.partials, s,
                       G.total_blocks, G.grid, G.prio | (G.tie_cap << 8));
}

template <uint32_t A, bool BITS>
void launch_ising_fcc_tile(const Group& G, hipStream_t st, uint32_t s) {
    switch (G.tile) {
        case 1: launch_ising_fcc<A, 1, BITS>(G, st, s); break;
        case 2: launch_ising_fcc<A, 2, BITS>(G, st, s); break;
        case 8: launch_ising_fcc<A, 8, BITS>(G, st, s); break;
        default: launch_ising_fcc<A, 4, BITS>(G, st, s); break;
    }
}

// kind: 0 Ising int8, 1 Ising bit-packed, 2 Heisenberg.
template <uint32_t A>
void launch_fcc_colour(const Group& G, hipStream_t st, uint32_t s, int kind) {
    const dim3 grid(G.grid), block(vg::kBlock);
    if (kind == 0)
        launch_ising_fcc_tile<A, false>(G, st, s);
    else if (kind == 1)
        launch_ising_fcc_tile<A, true>(G, st, s);
    else
        hipLaunchKernelGGL(vg::heis_fcc_pass<A>, grid, block, 0, st, G.d_chains, G.d_block_chain,
                           (float*)G.arena, (int64_t*)G.partials, s, G.total_blocks, G.grid);
}

void launch_fcc_pass(const Group& G, hipStream_t st, uint32_t c, uint32_t s, int kind) {
    switch (c) {
        case 0: launch_fcc_colour<0>(G, st, s, kind); break;
        case 1: launch_fcc_colour<1>(G, st, s, kind); break;
        case 2: launch_fcc_colour<2>(G, st, s, kind); break;
        default: launch_fcc_colour<3>(G, st, s, kind); break;
    }
}

// One colour pass of a z-marching Ising group (sc, bcc, fcc; int8 or bits).
void launch_zmarch_pass(const Group& G, int g, hipStream_t st, uint32_t c, uint32_t s) {
    switch (g) {
        case kIsingBits: launch_ising_pass<true>(G, st, c, s); break;
        case kIsingBcc: launch_bcc_pass<false>(G, st, c, s); break;
        case kIsingBccBits: launch_bcc_pass<true>(G, st, c, s); break;
        case kIsingFcc: launch_fcc_pass(G, st, c, s, 0); break;
        case kIsingFccBits: launch_fcc_pass(G, st, c, s, 1); break;
        default: launch_ising_pass<false>(G, st, c, s); break;
    }
}

// Enqueues the sweeps of one block (nsw sweeps, every colour of every group,
// then the per-chain reductions) on the batch stream.
int enqueue_block(vg_batch* B, uint32_t nsw) {
    hipStream_t st = B->ctx->stream;
    const uint64_t nrec = (uint64_t)B->chains.size() * nsw * 5u;
    hipLaunchKernelGGL(vg::zero_i64, dim3((uint32_t)((nrec + 255) / 256)), dim3(256), 0, st,
                       B->d_records, nrec);
    for (uint32_t s = 0; s < nsw; ++s) {
        for (int g = 0; g < kNumGroups; ++g) {
            Group& G = B->groups[g];
            if (G.members.empty()) continue;
            for (uint32_t c = 0; c < passes_of(G, g); ++c) {
                const bool timed = B->timing && !B->use_graph;
                if (timed) {
                    const uint32_t k = B->pass_events_used;
                    if (k + 2 > B->pass_events.size()) return fail(VG_ERR_DEVICE, "timing events exhausted");
                    VG_HIP(hipEventRecord(B->pass_events[k], st));
                }
                const dim3 grid(G.grid), block(vg::kBlock);
                if (g == kIsingSc)
                    launch_ising_pass<false>(G, st, c, s);
                else if (g == kIsingBits)
                    launch_ising_pass<true>(G, st, c, s);
                else if (g == kIsingBcc)
                    launch_bcc_pass<false>(G, st, c, s);
                else if (g == kIsingBccBits)
                    launch_bcc_pass<true>(G, st, c, s);
                else if (g == kIsingFcc)
                    launch_fcc_pass(G, st, c, s, 0);
                else if (g == kIsingFccBits)
                    launch_fcc_pass(G, st, c, s, 1);
                else if (g == kHeisSc)
                    hipLaunchKernelGGL(vg::heis_sc_pass, grid, block, 0, st, G.d_chains,
                                       G.d_block_chain, (float*)G.arena, (int64_t*)G.partials,
                                       c, s, G.total_blocks, G.grid);
                else if (g == kHeisBcc)
                    hipLaunchKernelGGL(vg::heis_bcc_pass, grid, block, 0, st, G.d_chains,
                                       G.d_block_chain, (float*)G.arena, (int64_t*)G.partials,
                                       c, s, G.total_blocks, G.grid);
                else if (g == kIsingGraph)
                    hipLaunchKernelGGL(vg::graph_ising_pass, dim3(G.total_blocks), block, 0, st,
                                       G.d_chains, G.d_block_chain, (uint8_t*)G.arena,
                                       (int64_t*)G.partials, c, s, G.total_blocks, G.npass);
                else if (g == kHeisGraph)
                    hipLaunchKernelGGL(vg::graph_heis_pass, dim3(G.total_blocks), block, 0, st,
                                       G.d_chains, G.d_block_chain, (uint8_t*)G.arena,
                                       (int64_t*)G.partials, c, s, G.total_blocks, G.npass);
                else
                    launch_fcc_pass(G, st, c, s, 2);
                if (timed) {
                    VG_HIP(hipEventRecord(B->pass_events[B->pass_events_used + 1], st));
                    B->pass_events_used += 2;
                }
            }
        }
    }
    for (int g = 0; g < kNumGroups; ++g) {
        Group& G = B->groups[g];
        if (G.members.empty()) continue;
        dim3 grid((uint32_t)G.members.size(), nsw);
        if (g == kIsingWolff) {
            // chains of at most kWolffLdsSites sites in LDS, the rest in
            // global memory; each kernel's blocks skip the other's chains
            uint32_t lds_words = 0;
            bool large = false;
            for (const ChainDev& d : G.host) {
                if (d.nsites <= vg::kWolffLdsSites)
                    lds_words = std::max(lds_words, 2u * (uint32_t)((d.nsites + 31) / 32) +
                                                        vg::kWolffLdsQueue);
                else
                    large = true;
            }
            const dim3 chains_grid((uint32_t)G.members.size());
            if (lds_words)
                hipLaunchKernelGGL(vg::wolff_run_lds, chains_grid, dim3(256), lds_words * 4u, st,
                                   G.d_chains, (uint8_t*)G.arena, nsw, B->d_records);
            if (large)
                hipLaunchKernelGGL(vg::wolff_run, chains_grid, dim3(256), 0, st, G.d_chains,
                                   (uint8_t*)G.arena, nsw, B->d_records);
        }
        else if (is_ising(g) && !is_graph(g))
            hipLaunchKernelGGL(vg::ising_reduce, grid, dim3(vg::kBlock), 0, st, G.d_chains,
                               (const int32_t*)G.partials, G.total_blocks, colours_of(g), nsw,
                               B->d_records);
        else
            hipLaunchKernelGGL(vg::heis_reduce, grid, dim3(vg::kBlock), 0, st, G.d_chains,
                               (const int64_t*)G.partials, G.total_blocks, passes_of(G, g), nsw,
                               B->d_records);
    }
    VG_HIP(hipGetLastError());
    return VG_OK;
}

// The captured graph of a block of nsw sweeps, captured, instantiated and
// uploaded to the device on first use.
int block_graph(vg_batch* B, uint32_t nsw, hipGraphExec_t* out) {
    auto it = B->graphs.find(nsw);
    if (it == B->graphs.end()) {
        hipStream_t st = B->ctx->stream;
        // Capture on an idle stream (an earlier block may still be queued).
        VG_HIP(hipStreamSynchronize(st));
        hipGraph_t graph = nullptr;
        VG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_block(B, nsw);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc != VG_OK) {
            if (graph) quiet(hipGraphDestroy(graph));
            return rc;
        }
        if (e != hipSuccess) return fail(VG_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        quiet(hipGraphDestroy(graph));
        if (e != hipSuccess)
            return fail(VG_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
        it = B->graphs.emplace(nsw, exec).first;
        VG_HIP(hipGraphUpload(exec, st));
    }
    *out = it->second;
    return VG_OK;
}

int launch_block(vg_batch* B, uint32_t nsw) {
    if (!B->use_graph) return enqueue_block(B, nsw);
    hipGraphExec_t exec = nullptr;
    int rc = block_graph(B, nsw, &exec);
    if (rc) return rc;
    VG_HIP(hipGraphLaunch(exec, B->ctx->stream));
    return VG_OK;
}

double sums_scale(int spin) { return spin == VG_SPIN_HEISENBERG ? 1.0 / 134217728.0 : 1.0; }

// Host side of the observables.  E follows Hamiltonian::total_energy of the
// reference (energy.rs:55-59 default for Compound; each term's own override
// for a bare term) and |M| follows State::magnetization().magnitude()
// (state.rs:86-92, 150-160, 291-296).
void observables(const ChainHost& ch, const vg_sweep_sums& r, double* E, double* M) {
    const vg_hamiltonian& H = ch.desc.hamiltonian;
    const double sc = sums_scale(ch.desc.spin);
    const double N = (double)ch.n_sites;
    // graph chains: bond = sum over i != j of J_ij s_i.s_j (couplings inside)
    // at 2^-27, and sum_i e_i of the exchange is -(bond + sum_i J_ii)
    const bool graph = is_graph(ch.group);
    const double bond = (double)r.bond * (graph ? 1.0 / 134217728.0 : sc);
    double m[3] = {(double)r.mag[0] * sc, (double)r.mag[1] * sc, (double)r.mag[2] * sc};
    double axis2 = (double)r.axis2 * sc;
    double hdir[3] = {ch.field_dir[0], ch.field_dir[1], ch.field_dir[2]};
    if (ch.desc.spin != VG_SPIN_HEISENBERG) {
        axis2 = N;  // (s.axis)^2 = 1 for Ising spins
        hdir[0] = hdir[1] = 0.0;
        hdir[2] = ch.field_dir[2] >= 0.0 ? 1.0 : -1.0;
    }
    const double hm = std::fabs(ch.field_mag);
    const double mdoth = m[0] * hdir[0] + m[1] * hdir[1] + m[2] * hdir[2];
    double e = 0.0;
    if (H.compound) {
        if (H.terms & VG_TERM_EXCHANGE) e += graph ? -(bond + ch.diag) : -2.0 * H.exchange * bond;
        if (H.terms & VG_TERM_ZEEMAN) e += hm * mdoth;
        if (H.terms & VG_TERM_ANISOTROPY) e += H.anisotropy * axis2;
        if (H.terms & VG_TERM_GAUGE) e += H.gauge * N;
    } else if (H.terms & VG_TERM_EXCHANGE) {
        e = graph ? -(bond + ch.diag) / 2.0 : -H.exchange * bond;
    } else if (H.terms & VG_TERM_ZEEMAN) {
        e = -hm * mdoth;
    } else if (H.terms & VG_TERM_ANISOTROPY) {
        e = axis2;
    } else {
        e = H.gauge * N;
    }
    double mag;
    if (ch.desc.spin == VG_SPIN_HEISENBERG) {
        mag = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        if (std::fabs(mag) < DBL_EPSILON) mag = 0.0;
    } else {
        mag = std::fabs(m[2]);
    }
    if (E) *E = e;
    if (M) *M = mag;
}

// Runs `sweeps` sweeps of every chain in blocks of up to max_block_sweeps.
// Blocks are pipelined two deep: block k+1 (descriptors, graph or launches,
// record copy) is queued before the host waits for block k's records, so
// the device runs from block to block without waiting for the host.  The
// per-pass timing mode (plain launches with events per pass) runs one block
// at a time.
int run_impl(vg_batch* B, uint64_t sweeps, vg_sweep_sums* sums, double* energy,
             double* magnetization) {
    if (!B) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    const uint64_t nch = B->chains.size();
    // The sweep counter keys Philox with 56 bits (low word + 24 high bits).
    for (const ChainHost& ch : B->chains)
        if (sweeps > (1ull << 56) - ch.sweeps)
            return fail(VG_ERR_INVALID_ARGUMENT, "sweep counter would pass 2^56");
    hipStream_t st = B->ctx->stream;
    VG_HIP(hipSetDevice(B->ctx->device));
    B->last_ms = 0.0;
    const bool timed = B->timing && !B->use_graph;
    const size_t depth = timed ? 1 : 2;
    std::vector<uint64_t> start(nch);
    for (uint64_t i = 0; i < nch; ++i) start[i] = B->chains[i].sweeps;
    struct Block {
        uint32_t nsw;
        uint64_t first;  // sweeps of this run before the block
        int slot;
    };
    Block inflight[2];
    size_t n_inflight = 0, head = 0;
    uint64_t queued = 0;
    int slot = 0;
    while (queued < sweeps || n_inflight > 0) {
        if (queued < sweeps && n_inflight < depth) {
            const uint32_t nsw = (uint32_t)std::min<uint64_t>(B->max_block_sweeps, sweeps - queued);
            for (uint64_t i = 0; i < nch; ++i) {
                ChainHost& ch = B->chains[i];
                B->groups[ch.group].host[ch.index].sweep0 = start[i] + queued;
            }
            for (int g = 0; g < kNumGroups; ++g) {
                int rc = upload_descriptors(B, g, slot);
                if (rc) return rc;
            }
            if (timed) {
                size_t need = 0;  // a start/stop pair per colour pass
                for (int g = 0; g < kNumGroups; ++g)
                    if (!B->groups[g].members.empty())
                        need += (size_t)nsw * 2 * passes_of(B->groups[g], g);
                while (B->pass_events.size() < need) {
                    hipEvent_t e;
                    VG_HIP(hipEventCreate(&e));
                    B->pass_events.push_back(e);
                }
                B->pass_events_used = 0;
            }
            VG_HIP(hipEventRecord(slot ? B->ev2 : B->ev0, st));
            int rc = launch_block(B, nsw);
            if (rc) return rc;
            VG_HIP(hipEventRecord(slot ? B->ev3 : B->ev1, st));
            VG_HIP(hipMemcpyAsync(slot ? B->h_records2 : B->h_records, B->d_records,
                                  nch * nsw * 5u * sizeof(int64_t), hipMemcpyDeviceToHost, st));
            VG_HIP(hipEventRecord(B->done[slot], st));
            inflight[(head + n_inflight) % 2] = Block{nsw, queued, slot};
            ++n_inflight;
            queued += nsw;
            slot ^= 1;
            if (queued < sweeps && n_inflight < depth) continue;  // fill the pipeline
        }
        // The oldest block: wait for its records and turn them into results.
        const Block blk = inflight[head];
        head = (head + 1) % 2;
        --n_inflight;
        VG_HIP(hipEventSynchronize(B->done[blk.slot]));
        float ms = 0.0f;
        VG_HIP(hipEventElapsedTime(&ms, blk.slot ? B->ev2 : B->ev0, blk.slot ? B->ev3 : B->ev1));
        B->last_ms += ms;
        const uint32_t nsw = blk.nsw;
        if (timed) {
            // Events were recorded in launch order: sweep, group, colour.
            uint32_t k = 0;
            for (uint32_t sw = 0; sw < nsw; ++sw)
                for (int g = 0; g < kNumGroups; ++g) {
                    if (B->groups[g].members.empty()) continue;
                    for (uint32_t c = 0; c < passes_of(B->groups[g], g); ++c, k += 2) {
                        float t = 0.0f;
                        VG_HIP(hipEventElapsedTime(&t, B->pass_events[k], B->pass_events[k + 1]));
                        B->pass_ms[g] += t;
                        B->pass_launches[g] += 1;
                    }
                }
        }
        const int64_t* rec = blk.slot ? B->h_records2 : B->h_records;
        for (uint64_t i = 0; i < nch; ++i) {
            ChainHost& ch = B->chains[i];
            for (uint32_t s = 0; s < nsw; ++s) {
                const int64_t* r = rec + (i * nsw + s) * 5u;
                vg_sweep_sums v;
                v.bond = r[0];
                v.mag[0] = r[1];
                v.mag[1] = r[2];
                v.mag[2] = r[3];
                v.axis2 = ch.desc.spin == VG_SPIN_HEISENBERG ? r[4] : (int64_t)ch.n_sites;
                const uint64_t o = i * sweeps + blk.first + s;
                if (sums) sums[o] = v;
                if (energy || magnetization)
                    observables(ch, v, energy ? energy + o : nullptr,
                                magnetization ? magnetization + o : nullptr);
            }
            ch.sweeps = start[i] + blk.first + nsw;
        }
    }
    return VG_OK;
}

int check_chain(const vg_batch* B, uint32_t chain) {
    if (!B) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    if (chain >= B->chains.size())
        return fail(VG_ERR_INVALID_ARGUMENT, "chain %u out of range (%zu chains)", chain,
                    B->chains.size());
    return VG_OK;
}

}  // namespace vgi

using namespace vgi;

extern "C" {

const char* vg_last_error(void) { return g_error.c_str(); }

const char* vg_version(void) { return "vegas-hip 0.1.0 (gfx950)"; }

double vg_sums_scale(int32_t spin) { return sums_scale(spin); }

void vg_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    const vg::U4 r = vg::philox(vg::U4{counter[0], counter[1], counter[2], counter[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

int vg_device_count(int* count) {
    if (!count) return fail(VG_ERR_INVALID_ARGUMENT, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return VG_OK;
}

int vg_ctx_create(int device, vg_ctx** out) {
    if (!out) return fail(VG_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(VG_ERR_NO_DEVICE, "no HIP device is visible");
    if (device < 0 || device >= n)
        return fail(VG_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, n);
    VG_HIP(hipSetDevice(device));
    vg_ctx* c = new vg_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(VG_ERR_DEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return VG_OK;
}

int vg_ctx_destroy(vg_ctx* ctx) {
    if (!ctx) return VG_OK;
    quiet(hipSetDevice(ctx->device));
    quiet(hipStreamSynchronize(ctx->stream));
    quiet(hipStreamDestroy(ctx->stream));
    delete ctx;
    return VG_OK;
}

}  // extern "C"

namespace vgi {

// Copies a host array into the arena on the batch stream.
template <class T>
int upload_section(vg_batch* B, Group& G, uint64_t off, const std::vector<T>& v) {
    if (v.empty()) return VG_OK;
    VG_HIP(hipMemcpyAsync((uint8_t*)G.arena + off, v.data(), v.size() * sizeof(T),
                          hipMemcpyHostToDevice, B->ctx->stream));
    return VG_OK;
}

// The colour classes, merged rows and site ids of graph chain k of group g.
int upload_graph(vg_batch* B, int g, size_t k, const GraphBuild& gr) {
    Group& G = B->groups[g];
    const ChainDev& d = G.host[k];
    int rc = upload_section(B, G, d.off[1], gr.ptr);
    if (!rc) rc = upload_section(B, G, d.off[2], gr.nbr);
    if (!rc) rc = upload_section(B, G, d.goff[1], gr.site_of_pos);
    if (!rc) rc = upload_section(B, G, d.goff[2], gr.cstart);
    if (rc) return rc;
    if (g == kIsingGraph) {
        rc = upload_section(B, G, d.off[3], gr.a);
        if (!rc) rc = upload_section(B, G, d.goff[0], gr.w);
    } else {  // Heisenberg couplings in fp32
        std::vector<float> af(gr.a.begin(), gr.a.end()), wf(gr.w.begin(), gr.w.end());
        rc = upload_section(B, G, d.off[3], af);
        if (!rc) rc = upload_section(B, G, d.goff[0], wf);
    }
    // the host arrays may go away once this returns
    const hipError_t e = hipStreamSynchronize(B->ctx->stream);
    if (!rc && e != hipSuccess) rc = fail(VG_ERR_DEVICE, "uploading a graph: %s", hipGetErrorString(e));
    return rc;
}

int create_impl(vg_ctx* ctx, uint32_t n_chains, const vg_chain_desc* descs, const vg_csr* graphs,
                vg_batch** out) {
    if (!ctx || !descs || !out) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (n_chains == 0) return fail(VG_ERR_INVALID_ARGUMENT, "a batch needs at least one chain");
    std::vector<GraphBuild> gb(n_chains);
    for (uint32_t i = 0; i < n_chains; ++i) {
        int rc = validate_desc(descs[i]);
        if (rc) return fail(rc, "chain %u: %s", i, g_error.c_str());
        if (descs[i].lattice.kind == VG_LATTICE_GRAPH) {
            if (!graphs)
                return fail(VG_ERR_INVALID_ARGUMENT,
                            "chain %u has an exchange matrix: create the batch with "
                            "vg_batch_create_csr", i);
            rc = build_graph(graphs[i], gb[i]);
            if (rc) return fail(rc, "chain %u: %s", i, g_error.c_str());
        }
    }
    VG_HIP(hipSetDevice(ctx->device));
    vg_batch* B = new vg_batch;
    B->ctx = ctx;
    B->chains.resize(n_chains);
    for (uint32_t i = 0; i < n_chains; ++i) {
        ChainHost& ch = B->chains[i];
        ch.desc = descs[i];
        const vg_lattice& L = ch.desc.lattice;
        ch.group = group_of(ch.desc);
        if (is_graph(ch.group)) {
            ch.n_sites = graphs[i].n_sites;
            ch.diag = gb[i].diag;
            ch.graph_pos = gb[i].pos_of_site;
        } else {
            ch.n_sites = sites_of(L);
        }
        Group& G = B->groups[ch.group];
        ch.index = (uint32_t)G.members.size();
        G.members.push_back(i);
    }
    // Lay out every group: arena offsets, work blocks, partial buffers.
    for (int g = 0; g < kNumGroups; ++g) {
        Group& G = B->groups[g];
        if (G.members.empty()) continue;
        G.host.resize(G.members.size());
        uint64_t offset = 0;  // in elements (bytes for Ising, floats for Heisenberg)
        for (size_t k = 0; k < G.members.size(); ++k) {
            ChainHost& ch = B->chains[G.members[k]];
            ChainDev& d = G.host[k];
            int rc = init_chain_dev(ch, g, G.members[k], d);
            if (rc) {
                free_batch(B);
                return rc;
            }
            const uint64_t elem_align = 256 / elem_bytes(g);  // 256-byte chain alignment
            if (is_graph(g)) {  // bytes: spins, row offsets, neighbours, a, w, site ids, classes
                const GraphBuild& gr = gb[G.members[k]];
                const uint64_t n = ch.n_sites, m = gr.nbr.size();
                const uint64_t vb = g == kIsingGraph ? 8 : 4;
                const uint64_t sizes[7] = {g == kIsingGraph ? n : 12 * n, 4 * (n + 1), 4 * m,
                                           vb * m, vb * m, 4 * n, 4 * ((uint64_t)gr.ncol + 1)};
                uint64_t offs[7];
                for (int part = 0; part < 7; ++part) {
                    offset = round_up(offset, 256);
                    offs[part] = offset;
                    offset += sizes[part];
                }
                for (int part = 0; part < 4; ++part) d.off[part] = offs[part];
                for (int part = 0; part < 3; ++part) d.goff[part] = offs[4 + part];
                d.ncol = gr.ncol;
                d.items = gr.max_class;
                G.npass = std::max(G.npass, gr.ncol);
                continue;
            }
            if (g == kIsingWolff) {  // bytes: spins, stamps, queue
                const uint64_t sizes[3] = {d.nsites, 4 * d.nsites, 4 * d.nsites};
                for (int part = 0; part < 3; ++part) {
                    offset = round_up(offset, 256);
                    d.off[part] = offset;
                    offset += sizes[part];
                }
                continue;
            }
            for (uint32_t c = 0; c < d.ncol; ++c) {
                offset = round_up(offset, elem_align);
                d.off[c] = offset;
                offset += colour_elems(g, d);
            }
            if (g == kIsingSc) {  // bit-packed shadows of the two planes
                d.spitch = (uint32_t)round_up((d.w + 31) / 32 * 4, 16);
                for (uint32_t c = 0; c < 2; ++c) {
                    offset = round_up(offset, elem_align);
                    d.soff[c] = offset;
                    offset += (uint64_t)d.rows * d.spitch;
                }
                d.shadow = 1;
                G.shadow = true;
            }
        }
        G.arena_bytes = std::max<uint64_t>(offset * elem_bytes(g), 256);
        hipError_t e = hipMalloc(&G.arena, G.arena_bytes);
        if (e == hipSuccess) e = hipMemset(G.arena, 0, G.arena_bytes);
        if (e == hipSuccess) e = hipMalloc(&G.d_chains, G.host.size() * sizeof(ChainDev));
        if (e == hipSuccess) e = hipHostMalloc(&G.h_staging, G.host.size() * sizeof(ChainDev));
        if (e == hipSuccess) e = hipHostMalloc(&G.h_staging2, G.host.size() * sizeof(ChainDev));
        G.chains_cap = (uint32_t)G.host.size();
        if (e != hipSuccess) {
            free_batch(B);
            return fail(e == hipErrorOutOfMemory ? VG_ERR_OUT_OF_MEMORY : VG_ERR_DEVICE,
                        "allocating the chain arena (%llu bytes): %s",
                        (unsigned long long)G.arena_bytes, hipGetErrorString(e));
        }
        if (g == kIsingWolff) {
            // the LDS Wolff kernel's dynamic LDS goes past the 64 KiB default
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(vg::wolff_run_lds),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(vg::kWolffLdsWordsMax * 4u));
            if (e != hipSuccess) {
                free_batch(B);
                return fail(VG_ERR_DEVICE, "Wolff LDS size: %s", hipGetErrorString(e));
            }
        }
        if (is_graph(g)) {
            for (size_t k = 0; k < G.members.size(); ++k) {
                int rc = upload_graph(B, g, k, gb[G.members[k]]);
                if (rc) {
                    free_batch(B);
                    return rc;
                }
            }
        }
        // Work layout with the finest tile sizes the partial buffers.
        int rc = layout_work(B, g, 1);
        if (rc) {
            free_batch(B);
            return rc;
        }
    }
    // Sweep-block length: keep the partial buffers within 64 MiB per group.
    uint64_t max_per_sweep = 1;
    for (Group& G : B->groups) max_per_sweep = std::max(max_per_sweep, G.partial_bytes_per_sweep);
    B->max_block_sweeps =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(256, (64ull << 20) / max_per_sweep));
    for (int g = 0; g < kNumGroups; ++g) {
        Group& G = B->groups[g];
        if (G.members.empty()) continue;
        hipError_t e = hipMalloc(&G.partials, G.partial_bytes_per_sweep * B->max_block_sweeps);
        if (e != hipSuccess) {
            free_batch(B);
            return fail(VG_ERR_OUT_OF_MEMORY, "allocating partial sums: %s", hipGetErrorString(e));
        }
        if (is_zmarch(g)) {
            int rc = layout_work(B, g, kDefaultTile);
            if (rc) {
                free_batch(B);
                return rc;
            }
        }
    }
    const uint64_t rec_bytes = (uint64_t)n_chains * B->max_block_sweeps * 5u * sizeof(int64_t);
    hipError_t e = hipMalloc(&B->d_records, rec_bytes);
    if (e == hipSuccess) e = hipHostMalloc(&B->h_records, rec_bytes);
    if (e == hipSuccess) e = hipHostMalloc(&B->h_records2, rec_bytes);
    for (hipEvent_t* ev : {&B->ev0, &B->ev1, &B->ev2, &B->ev3, &B->done[0], &B->done[1]})
        if (e == hipSuccess) e = hipEventCreate(ev);
    if (e != hipSuccess) {
        free_batch(B);
        return fail(VG_ERR_DEVICE, "allocating records: %s", hipGetErrorString(e));
    }
    for (int g = 0; g < kNumGroups; ++g) {
        int rc = upload_descriptors(B, g);
        if (rc) {
            free_batch(B);
            return rc;
        }
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_batch(B);
        return fail(VG_ERR_DEVICE, "%s", hipGetErrorString(e));
    }
    *out = B;
    return VG_OK;
}

}  // namespace vgi

extern "C" {

int vg_batch_create(vg_ctx* ctx, uint32_t n_chains, const vg_chain_desc* descs, vg_batch** out) {
    return create_impl(ctx, n_chains, descs, nullptr, out);
}

int vg_batch_create_csr(vg_ctx* ctx, uint32_t n_chains, const vg_chain_desc* descs,
                        const vg_csr* exchange, vg_batch** out) {
    if (!exchange) return fail(VG_ERR_INVALID_ARGUMENT, "null exchange matrices");
    return create_impl(ctx, n_chains, descs, exchange, out);
}

int vg_batch_colours(const vg_batch* batch, uint32_t chain, uint32_t* colours) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!colours) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    const ChainHost& ch = batch->chains[chain];
    *colours = is_graph(ch.group) ? batch->groups[ch.group].host[ch.index].ncol
                                  : colours_of(ch.group);
    return VG_OK;
}

int vg_batch_sums_scales(const vg_batch* batch, uint32_t chain, double scales[5]) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!scales) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    const ChainHost& ch = batch->chains[chain];
    for (int k = 0; k < 5; ++k) scales[k] = sums_scale(ch.desc.spin);
    if (is_graph(ch.group)) scales[0] = 1.0 / 134217728.0;
    return VG_OK;
}

int vg_batch_destroy(vg_batch* batch) {
    if (!batch) return VG_OK;
    quiet(hipSetDevice(batch->ctx->device));
    quiet(hipStreamSynchronize(batch->ctx->stream));
    return free_batch(batch);
}

int vg_batch_size(const vg_batch* batch, uint32_t* n) {
    if (!batch || !n) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *n = (uint32_t)batch->chains.size();
    return VG_OK;
}

int vg_batch_num_sites(const vg_batch* batch, uint32_t chain, uint64_t* n) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!n) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *n = batch->chains[chain].n_sites;
    return VG_OK;
}

int vg_batch_set_thermostat(vg_batch* batch, uint32_t chain, double temperature,
                            double field_magnitude, const double field_orientation[3]) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (std::isnan(temperature) || std::isnan(field_magnitude))
        return fail(VG_ERR_INVALID_ARGUMENT, "NaN thermostat parameter");
    ChainHost& ch = batch->chains[chain];
    ch.temperature = temperature < DBL_EPSILON ? DBL_EPSILON : temperature;
    ch.field_mag = field_magnitude;
    if (field_orientation) {
        for (int k = 0; k < 3; ++k) ch.field_dir[k] = field_orientation[k];
    } else {
        ch.field_dir[0] = ch.field_dir[1] = 0.0;
        ch.field_dir[2] = 1.0;
    }
    refresh_params(batch, chain);
    return VG_OK;
}

int vg_batch_randomize(vg_batch* batch) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    hipStream_t st = batch->ctx->stream;
    VG_HIP(hipSetDevice(batch->ctx->device));
    for (int g = 0; g < kNumGroups; ++g) {
        int rc = upload_descriptors(batch, g);
        if (rc) return rc;
    }
    for (ChainHost& ch : batch->chains) {
        Group& G = batch->groups[ch.group];
        const ChainDev& d = G.host[ch.index];
        const ChainDev* dp = G.d_chains + ch.index;
        if (ch.group == kIsingWolff) {
            hipLaunchKernelGGL(vg::wolff_init, dim3((uint32_t)((d.nsites + 255) / 256)), dim3(256),
                               0, st, dp, (uint8_t*)G.arena);
        } else if (ch.group == kIsingGraph) {
            hipLaunchKernelGGL(vg::graph_ising_init, dim3((uint32_t)((d.nsites + 255) / 256)),
                               dim3(256), 0, st, dp, (uint8_t*)G.arena);
        } else if (ch.group == kHeisGraph) {
            hipLaunchKernelGGL(vg::graph_heis_init, dim3((uint32_t)((d.nsites + 255) / 256)),
                               dim3(256), 0, st, dp, (uint8_t*)G.arena);
        } else if (is_ising(ch.group)) {
            launch_ising_init(ch.group, G, d, dp, st);
        } else {
            const uint64_t n = (uint64_t)d.ncol * d.rows * d.w;
            hipLaunchKernelGGL(vg::heis_init, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                               dp, (float*)G.arena);
        }
        ch.sweeps = 0;
    }
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(st));
    return VG_OK;
}

int vg_batch_set_ising(vg_batch* batch, uint32_t chain, const int8_t* spins) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!spins) return fail(VG_ERR_INVALID_ARGUMENT, "null spins");
    ChainHost& ch = batch->chains[chain];
    if (!is_ising(ch.group)) return fail(VG_ERR_INVALID_ARGUMENT, "chain %u is not Ising", chain);
    for (uint64_t n = 0; n < ch.n_sites; ++n)
        if (spins[n] != 1 && spins[n] != -1)
            return fail(VG_ERR_INVALID_ARGUMENT, "spin %llu is %d, not +1 or -1",
                        (unsigned long long)n, (int)spins[n]);
    Group& G = batch->groups[ch.group];
    hipStream_t st = batch->ctx->stream;
    VG_HIP(hipSetDevice(batch->ctx->device));
    rc = upload_descriptors(batch, ch.group);
    if (rc) return rc;
    if (ch.group == kIsingWolff) {  // natural order, int8 +-1
        VG_HIP(hipMemcpyAsync((uint8_t*)G.arena + G.host[ch.index].off[0], spins, ch.n_sites,
                              hipMemcpyHostToDevice, st));
        VG_HIP(hipStreamSynchronize(st));
        return VG_OK;
    }
    if (ch.group == kIsingGraph) {  // colour-sorted positions, 1 = Up
        std::vector<uint8_t> v(ch.n_sites);
        for (uint64_t i = 0; i < ch.n_sites; ++i) v[ch.graph_pos[i]] = spins[i] > 0 ? 1 : 0;
        VG_HIP(hipMemcpyAsync((uint8_t*)G.arena + G.host[ch.index].off[0], v.data(), ch.n_sites,
                              hipMemcpyHostToDevice, st));
        VG_HIP(hipStreamSynchronize(st));
        return VG_OK;
    }
    int8_t* tmp = nullptr;
    VG_HIP(hipMalloc(&tmp, ch.n_sites));
    hipError_t e = hipMemcpyAsync(tmp, spins, ch.n_sites, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        if (is_bits(ch.group)) {
            const uint64_t nb = (uint64_t)G.host[ch.index].ncol * G.host[ch.index].plane;
            hipLaunchKernelGGL(vg::ising_scatter_bits, dim3((uint32_t)((nb + 255) / 256)),
                               dim3(256), 0, st, G.d_chains + ch.index, (uint8_t*)G.arena, tmp);
        } else {
            hipLaunchKernelGGL(vg::ising_scatter, dim3((uint32_t)((ch.n_sites + 255) / 256)),
                               dim3(256), 0, st, G.d_chains + ch.index, (uint8_t*)G.arena, tmp,
                               ch.n_sites);
            if (ch.group == kIsingSc)
                launch_shadow_build(G, G.host[ch.index], G.d_chains + ch.index, st);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    quiet(hipFree(tmp));
    if (e != hipSuccess) return fail(VG_ERR_DEVICE, "uploading spins: %s", hipGetErrorString(e));
    return VG_OK;
}

int vg_batch_get_ising(vg_batch* batch, uint32_t chain, int8_t* spins) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!spins) return fail(VG_ERR_INVALID_ARGUMENT, "null spins");
    ChainHost& ch = batch->chains[chain];
    if (!is_ising(ch.group)) return fail(VG_ERR_INVALID_ARGUMENT, "chain %u is not Ising", chain);
    Group& G = batch->groups[ch.group];
    hipStream_t st = batch->ctx->stream;
    VG_HIP(hipSetDevice(batch->ctx->device));
    rc = upload_descriptors(batch, ch.group);
    if (rc) return rc;
    if (ch.group == kIsingWolff) {
        VG_HIP(hipMemcpyAsync(spins, (const uint8_t*)G.arena + G.host[ch.index].off[0],
                              ch.n_sites, hipMemcpyDeviceToHost, st));
        VG_HIP(hipStreamSynchronize(st));
        return VG_OK;
    }
    if (ch.group == kIsingGraph) {
        std::vector<uint8_t> v(ch.n_sites);
        VG_HIP(hipMemcpyAsync(v.data(), (const uint8_t*)G.arena + G.host[ch.index].off[0],
                              ch.n_sites, hipMemcpyDeviceToHost, st));
        VG_HIP(hipStreamSynchronize(st));
        for (uint64_t i = 0; i < ch.n_sites; ++i) spins[i] = v[ch.graph_pos[i]] ? 1 : -1;
        return VG_OK;
    }
    int8_t* tmp = nullptr;
    VG_HIP(hipMalloc(&tmp, ch.n_sites));
    if (is_bits(ch.group))
        hipLaunchKernelGGL(vg::ising_gather_bits, dim3((uint32_t)((ch.n_sites + 255) / 256)),
                           dim3(256), 0, st, G.d_chains + ch.index, (const uint8_t*)G.arena, tmp,
                           ch.n_sites);
    else
        hipLaunchKernelGGL(vg::ising_gather, dim3((uint32_t)((ch.n_sites + 255) / 256)),
                           dim3(256), 0, st, G.d_chains + ch.index, (const uint8_t*)G.arena, tmp,
                           ch.n_sites);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(spins, tmp, ch.n_sites, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    quiet(hipFree(tmp));
    if (e != hipSuccess) return fail(VG_ERR_DEVICE, "downloading spins: %s", hipGetErrorString(e));
    return VG_OK;
}

int vg_batch_set_heisenberg(vg_batch* batch, uint32_t chain, const double* xyz) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!xyz) return fail(VG_ERR_INVALID_ARGUMENT, "null spins");
    ChainHost& ch = batch->chains[chain];
    if (is_ising(ch.group))
        return fail(VG_ERR_INVALID_ARGUMENT, "chain %u is not Heisenberg", chain);
    Group& G = batch->groups[ch.group];
    hipStream_t st = batch->ctx->stream;
    VG_HIP(hipSetDevice(batch->ctx->device));
    rc = upload_descriptors(batch, ch.group);
    if (rc) return rc;
    if (ch.group == kHeisGraph) {  // x, y, z planes over the colour-sorted positions
        const uint64_t n = ch.n_sites;
        std::vector<float> v(3 * n);
        for (uint64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) v[a * n + ch.graph_pos[i]] = (float)xyz[3 * i + a];
        VG_HIP(hipMemcpyAsync((uint8_t*)G.arena + G.host[ch.index].off[0], v.data(),
                              v.size() * sizeof(float), hipMemcpyHostToDevice, st));
        VG_HIP(hipStreamSynchronize(st));
        return VG_OK;
    }
    double* tmp = nullptr;
    const uint64_t bytes = ch.n_sites * 3 * sizeof(double);
    VG_HIP(hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpyAsync(tmp, xyz, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(vg::heis_scatter, dim3((uint32_t)((ch.n_sites + 255) / 256)), dim3(256),
                           0, st, G.d_chains + ch.index, (float*)G.arena, tmp, ch.n_sites);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    quiet(hipFree(tmp));
    if (e != hipSuccess) return fail(VG_ERR_DEVICE, "uploading spins: %s", hipGetErrorString(e));
    return VG_OK;
}

int vg_batch_get_heisenberg(vg_batch* batch, uint32_t chain, double* xyz) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!xyz) return fail(VG_ERR_INVALID_ARGUMENT, "null spins");
    ChainHost& ch = batch->chains[chain];
    if (is_ising(ch.group))
        return fail(VG_ERR_INVALID_ARGUMENT, "chain %u is not Heisenberg", chain);
    Group& G = batch->groups[ch.group];
    hipStream_t st = batch->ctx->stream;
    VG_HIP(hipSetDevice(batch->ctx->device));
    rc = upload_descriptors(batch, ch.group);
    if (rc) return rc;
    if (ch.group == kHeisGraph) {
        const uint64_t n = ch.n_sites;
        std::vector<float> v(3 * n);
        VG_HIP(hipMemcpyAsync(v.data(), (const uint8_t*)G.arena + G.host[ch.index].off[0],
                              v.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        VG_HIP(hipStreamSynchronize(st));
        for (uint64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (double)v[a * n + ch.graph_pos[i]];
        return VG_OK;
    }
    double* tmp = nullptr;
    const uint64_t bytes = ch.n_sites * 3 * sizeof(double);
    VG_HIP(hipMalloc(&tmp, bytes));
    hipLaunchKernelGGL(vg::heis_gather, dim3((uint32_t)((ch.n_sites + 255) / 256)), dim3(256), 0,
                       st, G.d_chains + ch.index, (const float*)G.arena, tmp, ch.n_sites);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(xyz, tmp, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    quiet(hipFree(tmp));
    if (e != hipSuccess) return fail(VG_ERR_DEVICE, "downloading spins: %s", hipGetErrorString(e));
    return VG_OK;
}

int vg_batch_sweep_count(const vg_batch* batch, uint32_t chain, uint64_t* sweeps) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (!sweeps) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *sweeps = batch->chains[chain].sweeps;
    return VG_OK;
}

int vg_batch_set_sweep_count(vg_batch* batch, uint32_t chain, uint64_t sweeps) {
    int rc = check_chain(batch, chain);
    if (rc) return rc;
    if (sweeps >> 56) return fail(VG_ERR_INVALID_ARGUMENT, "sweep counter beyond 2^56");
    batch->chains[chain].sweeps = sweeps;
    return VG_OK;
}

int vg_batch_run(vg_batch* batch, uint64_t sweeps, double* energy, double* magnetization) {
    return run_impl(batch, sweeps, nullptr, energy, magnetization);
}

int vg_batch_run_sums(vg_batch* batch, uint64_t sweeps, vg_sweep_sums* sums) {
    if (!sums) return fail(VG_ERR_INVALID_ARGUMENT, "null sums");
    return run_impl(batch, sweeps, sums, nullptr, nullptr);
}

int vg_batch_prepare(vg_batch* batch, uint64_t sweeps) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    if (!batch->use_graph || sweeps == 0) return VG_OK;
    VG_HIP(hipSetDevice(batch->ctx->device));
    const uint64_t full = batch->max_block_sweeps;
    hipGraphExec_t exec = nullptr;
    if (sweeps >= full) {
        int rc = block_graph(batch, (uint32_t)full, &exec);
        if (rc) return rc;
    }
    if (sweeps % full) {
        int rc = block_graph(batch, (uint32_t)(sweeps % full), &exec);
        if (rc) return rc;
    }
    VG_HIP(hipStreamSynchronize(batch->ctx->stream));
    return VG_OK;
}

int vg_batch_set_graph(vg_batch* batch, int enabled) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    batch->use_graph = enabled != 0;
    return VG_OK;
}

int vg_batch_set_timing(vg_batch* batch, int enabled) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    batch->timing = enabled != 0;
    for (int g = 0; g < kNumGroups; ++g) {
        batch->pass_ms[g] = 0.0;
        batch->pass_launches[g] = 0;
    }
    return VG_OK;
}

int vg_batch_pass_stats(const vg_batch* batch, int32_t spin, double* total_ms, uint64_t* launches) {
    if (!batch || !total_ms || !launches) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    // every group whose chains have this spin kind
    *total_ms = 0.0;
    *launches = 0;
    for (int g = 0; g < kNumGroups; ++g)
        if (is_ising(g) == (spin != VG_SPIN_HEISENBERG)) {
            *total_ms += batch->pass_ms[g];
            *launches += batch->pass_launches[g];
        }
    return VG_OK;
}

int vg_batch_set_tile(vg_batch* batch, uint32_t tile) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    if (tile != 1 && tile != 2 && tile != 4 && tile != 8)
        return fail(VG_ERR_INVALID_ARGUMENT, "tile must be 1, 2, 4 or 8 (got %u)", tile);
    VG_HIP(hipSetDevice(batch->ctx->device));
    for (int g : {kIsingSc, kIsingBits, kIsingBcc, kIsingBccBits, kIsingFcc, kIsingFccBits}) {
        if (batch->groups[g].members.empty()) continue;
        int rc = layout_work(batch, g, tile);
        if (rc) return rc;
    }
    return VG_OK;
}

// The group whose tile vg_batch_get_tile and vg_batch_autotune report: the
// first of the z-marching Ising groups (sc, bcc, fcc; int8, then bit-packed)
// that has chains.
static int tile_group(const vg_batch* batch) {
    for (int g : {kIsingSc, kIsingBits, kIsingBcc, kIsingBccBits, kIsingFcc, kIsingFccBits})
        if (!batch->groups[g].members.empty()) return g;
    return kIsingSc;
}

int vg_batch_get_tile(const vg_batch* batch, uint32_t* tile) {
    if (!batch || !tile) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *tile = batch->groups[tile_group(batch)].tile;
    return VG_OK;
}

int vg_batch_autotune(vg_batch* batch, uint32_t trial_sweeps, uint32_t* chosen) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    uint32_t tile_now = 0;
    vg_batch_get_tile(batch, &tile_now);
    bool any = false;
    for (int g : {kIsingSc, kIsingBits, kIsingBcc, kIsingBccBits, kIsingFcc, kIsingFccBits})
        any = any || !batch->groups[g].members.empty();
    if (!any || trial_sweeps == 0) {
        if (chosen) *chosen = tile_now;
        return VG_OK;
    }
    VG_HIP(hipSetDevice(batch->ctx->device));
    hipStream_t st = batch->ctx->stream;
    // The trials advance every chain of the batch; all arenas and sweep
    // counters are restored after each, so tuning changes no trajectory.
    std::vector<void*> backup(kNumGroups, nullptr);
    auto free_backup = [&]() {
        for (void*& p : backup)
            if (p) quiet(hipFree(p));
    };
    for (int g = 0; g < kNumGroups; ++g) {
        Group& G = batch->groups[g];
        if (G.members.empty()) continue;
        hipError_t e = hipMalloc(&backup[g], G.arena_bytes);
        if (e == hipSuccess)
            e = hipMemcpyAsync(backup[g], G.arena, G.arena_bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) {
            free_backup();
            return fail(VG_ERR_OUT_OF_MEMORY, "autotune backup: %s", hipGetErrorString(e));
        }
    }
    std::vector<uint64_t> counters(batch->chains.size());
    for (size_t i = 0; i < batch->chains.size(); ++i) counters[i] = batch->chains[i].sweeps;
    auto restore = [&]() -> int {
        for (size_t i = 0; i < batch->chains.size(); ++i) batch->chains[i].sweeps = counters[i];
        for (int g = 0; g < kNumGroups; ++g) {
            Group& G = batch->groups[g];
            if (G.members.empty()) continue;
            VG_HIP(hipMemcpyAsync(G.arena, backup[g], G.arena_bytes, hipMemcpyDeviceToDevice, st));
        }
        VG_HIP(hipStreamSynchronize(st));
        return VG_OK;
    };
    int rc = VG_OK;
    const bool use_graph = batch->use_graph;
    batch->use_graph = true;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(VG_ERR_DEVICE, "autotune backup: %s", hipGetErrorString(e));
    for (int g : {kIsingSc, kIsingBits, kIsingBcc, kIsingBccBits, kIsingFcc, kIsingFccBits}) {
        Group& G = batch->groups[g];
        if (G.members.empty() || rc != VG_OK) continue;
        const uint32_t start_tile = G.tile;
        uint32_t best = start_tile;
        const uint32_t tiles[4] = {1, 2, 4, 8};
        // Five rounds over the tiles, each tile's median kept: a drift of
        // the clocks during tuning touches every tile alike, and one fast
        // outlier does not pick a tile.
        constexpr int kRounds = 5;
        double ms[4][kRounds] = {};
        for (int rep = 0; rep < kRounds && rc == VG_OK; ++rep)
            for (uint32_t ti = 0; ti < 4 && rc == VG_OK; ++ti) {
                rc = layout_work(batch, g, tiles[ti]);
                if (rc == VG_OK) rc = run_impl(batch, 2, nullptr, nullptr, nullptr);
                if (rc == VG_OK) rc = restore();
                // Timed from graph replays captured beforehand (plain
                // launches of passes this short measure the launch rate).
                if (rc == VG_OK) rc = vg_batch_prepare(batch, trial_sweeps);
                if (rc == VG_OK) rc = run_impl(batch, trial_sweeps, nullptr, nullptr, nullptr);
                if (rc == VG_OK) ms[ti][rep] = batch->last_ms;
                if (rc == VG_OK) rc = restore();
            }
        if (rc == VG_OK) {
            double best_ms = 1e300;
            for (uint32_t ti = 0; ti < 4; ++ti) {
                std::sort(ms[ti], ms[ti] + kRounds);
                if (ms[ti][kRounds / 2] < best_ms) {
                    best_ms = ms[ti][kRounds / 2];
                    best = tiles[ti];
                }
            }
        }
        if (rc != VG_OK) {
            // A trial failed part way (graph capture, memory): put the spins
            // and sweep counters back as they were, best effort, keeping the
            // trial's error message for the caller.
            const std::string msg = g_error;
            restore();
            layout_work(batch, g, start_tile);
            g_error = msg;
            break;
        }
        rc = layout_work(batch, g, best);
    }
    free_backup();
    batch->use_graph = use_graph;
    if (rc == VG_OK && chosen) *chosen = batch->groups[tile_group(batch)].tile;
    return rc;
}

int vg_batch_set_tie_capacity(vg_batch* batch, uint32_t records) {
    if (!batch) return fail(VG_ERR_INVALID_ARGUMENT, "null batch");
    if (records > 128)
        return fail(VG_ERR_INVALID_ARGUMENT, "tie capacity must be 0..128 (got %u)", records);
    for (Group& G : batch->groups) G.tie_cap = records;
    // captured graphs hold the old kernel arguments
    for (auto& kv : batch->graphs) quiet(hipGraphExecDestroy(kv.second));
    batch->graphs.clear();
    return VG_OK;
}

int vg_batch_last_run_ms(const vg_batch* batch, double* ms) {
    if (!batch || !ms) return fail(VG_ERR_INVALID_ARGUMENT, "null argument");
    *ms = batch->last_ms;
    return VG_OK;
}

}  // extern "C"
