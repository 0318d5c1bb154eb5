// Internal interface of the engine shared by the C ABI (capi.hip) and the
// job server (server.hip): chain and group bookkeeping, descriptor upload,
// work layout, the sweep-block driver and the host side of the observables.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/vegas_hip.h"
#include "vg_device.hpp"

namespace vg {
// sc pass args: (chains, block -> chain map, arena, partials, colour, sweep,
// total blocks, grid, options, blocks per chain when uniform else 0).
template <uint32_t K, uint32_t MODE>
__global__ void ising_sc_pass(const ChainDev*, const uint32_t*, uint8_t*, int32_t*, uint32_t,
                              uint32_t, uint32_t, uint32_t, uint32_t, uint32_t);
// Packs the int8 sc planes of one chain into its shadows.
__global__ void ising_shadow_build(const ChainDev*, uint8_t*);
__global__ void ising_init_bits(const ChainDev*, uint8_t*);
__global__ void ising_scatter_bits(const ChainDev*, uint8_t*, const int8_t*);
__global__ void ising_gather_bits(const ChainDev*, const uint8_t*, int8_t*, uint64_t);
__global__ void ising_reduce(const ChainDev*, const int32_t*, uint32_t, uint32_t, uint32_t,
                             int64_t*);
__global__ void ising_init(const ChainDev*, uint8_t*);
__global__ void ising_scatter(const ChainDev*, uint8_t*, const int8_t*, uint64_t);
__global__ void ising_gather(const ChainDev*, const uint8_t*, int8_t*, uint64_t);
__global__ void heis_sc_pass(const ChainDev*, const uint32_t*, float*, int64_t*, uint32_t,
                             uint32_t, uint32_t, uint32_t);
template <uint32_t K, bool BITS>
__global__ void ising_bcc_pass(const ChainDev*, const uint32_t*, uint8_t*, int32_t*, uint32_t,
                               uint32_t, uint32_t, uint32_t, uint32_t);
__global__ void heis_bcc_pass(const ChainDev*, const uint32_t*, float*, int64_t*, uint32_t,
                              uint32_t, uint32_t, uint32_t);
__global__ void heis_reduce(const ChainDev*, const int64_t*, uint32_t, uint32_t, uint32_t,
                            int64_t*);
// fcc passes: one instantiation per colour A (sublattice), args (chains,
// block -> chain map, arena, partials, sweep, total blocks, grid[, wave
// priorities on]).
template <uint32_t A, uint32_t K, bool BITS>
__global__ void ising_fcc_pass(const ChainDev*, const uint32_t*, uint8_t*, int32_t*, uint32_t,
                               uint32_t, uint32_t, uint32_t);
template <uint32_t A>
__global__ void heis_fcc_pass(const ChainDev*, const uint32_t*, float*, int64_t*, uint32_t,
                              uint32_t, uint32_t);
__global__ void wolff_run(const ChainDev*, uint8_t*, uint32_t, int64_t*);
__global__ void wolff_run_lds(const ChainDev*, uint8_t*, uint32_t, int64_t*);
__global__ void wolff_init(const ChainDev*, uint8_t*);
__global__ void heis_clear(const ChainDev*, float*);
__global__ void heis_init(const ChainDev*, float*);
__global__ void heis_scatter(const ChainDev*, float*, const double*, uint64_t);
__global__ void heis_gather(const ChainDev*, const float*, double*, uint64_t);
__global__ void zero_i64(int64_t*, uint64_t);
// Graph chains (an arbitrary exchange matrix): one pass per colour class,
// args (chains, block -> chain map, arena, partials, colour, sweep, total
// blocks, passes per sweep); initial states.
__global__ void graph_ising_pass(const ChainDev*, const uint32_t*, uint8_t*, int64_t*, uint32_t,
                                 uint32_t, uint32_t, uint32_t);
__global__ void graph_heis_pass(const ChainDev*, const uint32_t*, uint8_t*, int64_t*, uint32_t,
                                uint32_t, uint32_t, uint32_t);
__global__ void graph_ising_init(const ChainDev*, uint8_t*);
__global__ void graph_heis_init(const ChainDev*, uint8_t*);
}  // namespace vg

using vg::ChainDev;

namespace vgi {

// Thread-local message of the last failure (vg_last_error) and a setter
// that returns the status code.
extern thread_local std::string g_error;
int fail(int code, const char* fmt, ...);

#define VG_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(e_ == hipErrorOutOfMemory ? VG_ERR_OUT_OF_MEMORY : VG_ERR_DEVICE, \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,  \
                        __LINE__);                                                        \
    } while (0)

inline uint64_t round_up(uint64_t v, uint64_t m) { return (v + m - 1) / m * m; }

// Best-effort cleanup: errors there cannot be reported to anyone.
inline void quiet(hipError_t) {}

// Kernel groups: chains of one group share an arena and one pass kernel.
enum GroupKind {
    kIsingSc = 0,
    kHeisSc = 1,
    kIsingBcc = 2,
    kHeisBcc = 3,
    kIsingBits = 4,  // sc, 1 bit per site
    kIsingFcc = 5,
    kHeisFcc = 6,
    kIsingWolff = 7,  // Wolff cluster steps, any lattice, natural site order
    kIsingBccBits = 8,  // bcc, 1 bit per site
    kIsingFccBits = 9,  // fcc, 1 bit per site
    kIsingGraph = 10,   // arbitrary exchange matrix, int8 Ising, colour classes
    kHeisGraph = 11,    // arbitrary exchange matrix, fp32 Heisenberg
    kNumGroups = 12
};

inline bool is_ising(int g) {
    return g == kIsingSc || g == kIsingBcc || g == kIsingBits || g == kIsingFcc ||
           g == kIsingWolff || g == kIsingBccBits || g == kIsingFccBits || g == kIsingGraph;
}
// Groups of chains whose exchange is a matrix (VG_LATTICE_GRAPH).
inline bool is_graph(int g) { return g == kIsingGraph || g == kHeisGraph; }
// Groups whose planes hold one bit per site.
inline bool is_bits(int g) {
    return g == kIsingBits || g == kIsingBccBits || g == kIsingFccBits;
}
// Groups whose pass marches a tile of z planes per work item.
inline bool is_zmarch(int g) {
    return g == kIsingSc || g == kIsingBits || g == kIsingBcc || g == kIsingFcc ||
           g == kIsingBccBits || g == kIsingFccBits;
}
// Colours (planes, passes per sweep) of a lattice group's chains; Wolff
// chains have no colour passes; graph groups keep theirs in Group::npass.
inline uint32_t colours_of(int g) {
    return g == kIsingWolff || is_graph(g)
               ? 0u
               : (g == kIsingFcc || g == kHeisFcc || g == kIsingFccBits ? 4u : 2u);
}

inline int group_of(const vg_chain_desc& d) {
    const int k = d.lattice.kind;
    if (k == VG_LATTICE_GRAPH) return d.spin == VG_SPIN_HEISENBERG ? kHeisGraph : kIsingGraph;
    if (d.proposal == VG_PROPOSAL_WOLFF) return kIsingWolff;
    if (d.spin == VG_SPIN_HEISENBERG)
        return k == VG_LATTICE_FCC ? kHeisFcc : (k == VG_LATTICE_BCC ? kHeisBcc : kHeisSc);
    if (d.spin == VG_SPIN_ISING_BITS)
        return k == VG_LATTICE_FCC ? kIsingFccBits : (k == VG_LATTICE_BCC ? kIsingBccBits : kIsingBits);
    return k == VG_LATTICE_FCC ? kIsingFcc : (k == VG_LATTICE_BCC ? kIsingBcc : kIsingSc);
}

// Sites of a lattice of lx*ly*lz cells: 1 (sc), 2 (bcc) or 4 (fcc) per cell.
inline uint64_t sites_of(const vg_lattice& L) {
    const uint64_t cells = (uint64_t)L.lx * L.ly * L.lz;
    return L.kind == VG_LATTICE_SC ? cells : (L.kind == VG_LATTICE_BCC ? 2 * cells : 4 * cells);
}

struct Group {
    std::vector<uint32_t> members;  // batch chain indices
    std::vector<ChainDev> host;     // descriptors, uploaded each sweep block
    ChainDev* d_chains = nullptr;
    ChainDev* h_staging = nullptr;  // pinned copy used for the upload
    ChainDev* h_staging2 = nullptr;  // second one, for the block after (batches)
    uint32_t* d_block_chain = nullptr;
    uint32_t block_chain_cap = 0;
    uint32_t tile = 4;  // Ising sc/bcc/fcc (int8, bits): z planes per work item (template parameter)
    uint32_t total_blocks = 0;
    uint32_t grid = 0;  // total_blocks rounded up to a multiple of 8
    // blocks per chain when every chain of the group has the same number (a
    // pass then finds a block's chain without the map), else 0
    uint32_t uniform_nb = 0;
    uint32_t prio = 0;  // z-marching Ising passes: every workgroup resident at once (wave priorities on)
    // z-marching Ising passes: tie records a wave may queue (each pass holds
    // at most its own queue: 128 records in the sc pass, 64 in bcc and fcc)
    uint32_t tie_cap = 128;
    uint32_t npass = 0;     // graph groups: passes per sweep (the most colours of a member)
    void* arena = nullptr;
    uint64_t arena_bytes = 0;
    void* partials = nullptr;  // [sweep][colour][block][quantity]
    uint64_t partial_bytes_per_sweep = 0;
    // Capacities of d_chains / h_staging (descriptors) and of partials
    // (bytes), for batches whose membership changes (the job server).
    uint32_t chains_cap = 0;
    uint64_t partial_cap = 0;
    bool shared_arena = false;  // arena is owned by someone else (the server's pool)
    // int8 sc chains with bit-packed plane shadows (batches; the server's
    // pooled chains have none)
    bool shadow = false;
};

struct ChainHost {
    vg_chain_desc desc;
    uint64_t n_sites = 0;
    uint64_t sweeps = 0;
    double temperature = 2.8;  // Thermostat::new(2.8, Field::zero()) of input.rs:274
    double field_mag = 0.0;
    double field_dir[3] = {0.0, 0.0, 1.0};
    int group = 0;
    uint32_t index = 0;  // index in its group
    // Graph chains: sum of the matrix's diagonal (a constant -J_ii in every
    // e_i) and the position of each site in the colour-sorted order.
    double diag = 0.0;
    std::vector<uint32_t> graph_pos;
};

// Passes per sweep of group g.
inline uint32_t passes_of(const Group& G, int g) { return is_graph(g) ? G.npass : colours_of(g); }

}  // namespace vgi

struct vg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct vg_batch {
    vg_ctx* ctx = nullptr;
    std::vector<vgi::ChainHost> chains;
    vgi::Group groups[vgi::kNumGroups];
    uint32_t max_block_sweeps = 256;
    int64_t* d_records = nullptr;  // [chain][sweep][5]
    int64_t* h_records = nullptr;  // pinned
    int64_t* h_records2 = nullptr;  // pinned, for the block after (batches)
    bool use_graph = false;
    std::map<uint32_t, hipGraphExec_t> graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev2 = nullptr, ev3 = nullptr;  // the same, for the block after
    hipEvent_t done[2] = {nullptr, nullptr};  // records of a block are on the host
    double last_ms = 0.0;
    // Per-launch timing of the colour-pass kernels (plain launches only).
    bool timing = false;
    std::vector<hipEvent_t> pass_events;  // start/stop pairs
    uint32_t pass_events_used = 0;
    double pass_ms[vgi::kNumGroups] = {};
    uint64_t pass_launches[vgi::kNumGroups] = {};
};

namespace vgi {

// A graph chain's matrix prepared for the graph passes (graph_host.hip).
struct GraphBuild {
    uint32_t ncol = 0;                  // colour classes
    uint32_t max_class = 0;             // sites of the largest class
    std::vector<uint32_t> cstart;       // class starts, ncol + 1 (positions)
    std::vector<uint32_t> site_of_pos;  // positions sorted by (colour, site id)
    std::vector<uint32_t> pos_of_site;
    std::vector<uint32_t> ptr, nbr;     // merged rows by position; neighbours as positions
    std::vector<double> a, w;           // J_ij; J_ij + J_ji for lower-colour neighbours, else 0
    double diag = 0.0;                  // sum of the diagonal entries
};
int build_graph(const vg_csr& m, GraphBuild& out);

int validate_desc(const vg_chain_desc& d);
int coordination(int kind);
void refresh_params(vg_batch* B, uint32_t i);
// Acceptance tables (Ising) or fp32 parameters (Heisenberg) of a chain's
// current thermostat.
void ising_thresholds(const ChainHost& ch, ChainDev& d);
void heis_params(const ChainHost& ch, ChainDev& d);
int free_batch(vg_batch* B);
// Random initial state (State::rand_with_size) of an Ising chain of group g.
void launch_ising_init(int g, const Group& G, const ChainDev& d, const ChainDev* dp,
                       hipStream_t st);
uint32_t quantities(int group);
uint32_t quantity_bytes(int group);
// Copies group g's descriptors to the device through pinned staging buffer
// `slot` (0, or 1 for the second buffer of a batch).
int upload_descriptors(vg_batch* B, int g, int slot = 0);
int layout_work(vg_batch* B, int g, uint32_t tile);
// Whether `grid` workgroups of group g's z-marching pass at this tile are all
// resident at once (the falling wave priorities are used only then).
bool all_resident(int g, uint32_t tile, uint32_t grid);
void launch_zmarch_pass(const Group& G, int g, hipStream_t st, uint32_t c, uint32_t s);
int enqueue_block(vg_batch* B, uint32_t nsw);
int launch_block(vg_batch* B, uint32_t nsw);
double sums_scale(int spin);
void observables(const ChainHost& ch, const vg_sweep_sums& r, double* E, double* M);
// Geometry and parameters of chain `ch` as member `slot` of group g (sizes,
// magic divisors, acceptance tables); the plane offsets are set by the caller.
int init_chain_dev(const ChainHost& ch, int g, uint32_t slot, ChainDev& d);
// Elements (bytes for Ising, floats for Heisenberg) of one colour of a chain.
uint64_t colour_elems(int g, const ChainDev& d);
// Bytes of one element in group g's arena.
inline uint64_t elem_bytes(int g) { return is_ising(g) || is_graph(g) ? 1 : 4; }

}  // namespace vgi
