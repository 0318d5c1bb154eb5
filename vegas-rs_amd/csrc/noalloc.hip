// vg_run_sums_noalloc: the sweeps of vg_batch_run_sums for one Ising chain
// on caller-owned device buffers, allocating nothing.
//
// A batch allocates its arena, descriptors, block map, pass partials and
// records when it is created.  Here the caller brings the spins (in the
// batch arena's colour-plane layout) and the output sums; the descriptor,
// the block -> chain map and the pass partials live in `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs
// on holds until the process ends.  Host code reaches a device's copy only
// through the address hipGetSymbolAddress returns for it (kept per device):
// the array's name in host code is a host-side stand-in, not a device
// address.  The descriptor reaches the scratch through a kernel argument,
// so a call makes no host-to-device copy either.  Calls on one GPU share its
// scratch, and everything a call reads there it has written first, so
// nothing needs clearing; calls that overlap on one GPU race.
#include <mutex>

#include "engine.hpp"

namespace vg {

constexpr uint64_t kScratchBytes = 16ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr uint64_t kScratchDesc = 0;      // ChainDev
constexpr uint64_t kScratchMap = 4096;    // block -> chain map (all 0)
static_assert(sizeof(ChainDev) <= kScratchMap, "descriptor slot too small");

// Writes the chain's descriptor and the block -> chain map into the scratch.
__global__ void noalloc_setup(ChainDev d, ChainDev* __restrict__ out,
                              uint32_t* __restrict__ block_chain, uint32_t nblocks) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *out = d;
    if (i < nblocks) block_chain[i] = 0;
}

}  // namespace vg

namespace vgi {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch, looked up once per device.
int scratch_of(int device, uint8_t** out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(VG_ERR_INVALID_ARGUMENT, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(vg::scratch));
        if (e != hipSuccess || !p)
            return fail(VG_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_scratch[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch[device];
    return VG_OK;
}

}  // namespace
}  // namespace vgi

using namespace vgi;

extern "C" int vg_run_sums_noalloc(const vg_chain_desc* desc, double temperature,
                                   double field_magnitude, const double field_orientation[3],
                                   uint64_t sweep0, uint64_t sweeps, void* spins, int64_t* sums,
                                   void* stream) {
    if (!desc) return fail(VG_ERR_INVALID_ARGUMENT, "null descriptor");
    int rc = validate_desc(*desc);
    if (rc) return rc;
    const int g = group_of(*desc);
    if (!is_zmarch(g))
        return fail(VG_ERR_NOT_IMPLEMENTED,
                    "the allocation-free sweeps take Metropolis Ising chains (int8 or bit-packed) "
                    "on sc, bcc or fcc lattices");
    if (std::isnan(temperature) || std::isnan(field_magnitude))
        return fail(VG_ERR_INVALID_ARGUMENT, "NaN thermostat parameter");
    if (sweep0 >> 56 || sweeps > (1ull << 56) - sweep0)
        return fail(VG_ERR_INVALID_ARGUMENT, "sweep counter would pass 2^56");
    if (sweeps == 0) return VG_OK;
    if (!spins || !sums) return fail(VG_ERR_INVALID_ARGUMENT, "null spins or sums");

    // Thermostat and descriptor as a one-chain batch would hold them.
    ChainHost ch;
    ch.desc = *desc;
    ch.n_sites = sites_of(desc->lattice);
    ch.group = g;
    ch.temperature = temperature < DBL_EPSILON ? DBL_EPSILON : temperature;
    ch.field_mag = field_magnitude;
    if (field_orientation)
        for (int k = 0; k < 3; ++k) ch.field_dir[k] = field_orientation[k];
    ChainDev d;
    rc = init_chain_dev(ch, g, 0, d);
    if (rc) return rc;
    const uint64_t plane_bytes = round_up(d.plane, 256);
    for (uint32_t c = 0; c < d.ncol; ++c) d.off[c] = c * plane_bytes;
    const uint32_t tile = 4;
    d.nzc = (d.lz + tile - 1) / tile;
    const uint64_t items = (uint64_t)d.nzc * d.ly * d.segs;  // the sc pass's natural row order
    if (items > 0x7FFFFFFFull - vg::kBlock)
        return fail(VG_ERR_INVALID_ARGUMENT, "chain too large");
    d.items = (uint32_t)items;
    d.nblocks = (uint32_t)((items + vg::kBlock - 1) / vg::kBlock);
    d.block0 = 0;

    // Scratch layout: descriptor, block map, then as many sweeps' pass
    // partials as fit (8 int32 per workgroup, colour and sweep).
    const uint64_t map_end = vg::kScratchMap + 4ull * d.nblocks;
    const uint64_t part_off = round_up(map_end, 256);
    const uint64_t per_sweep = (uint64_t)d.ncol * d.nblocks * 8u * 4u;
    if (part_off + per_sweep > vg::kScratchBytes)
        return fail(VG_ERR_INVALID_ARGUMENT,
                    "chain of %llu sites needs %llu bytes of scratch per sweep; the scratch "
                    "holds %llu",
                    (unsigned long long)ch.n_sites, (unsigned long long)(part_off + per_sweep),
                    (unsigned long long)vg::kScratchBytes);
    const uint64_t chunk = std::min<uint64_t>((vg::kScratchBytes - part_off) / per_sweep, 4096);

    int device = 0;
    VG_HIP(hipGetDevice(&device));
    uint8_t* base = nullptr;
    rc = scratch_of(device, &base);
    if (rc) return rc;
    Group G;  // views of the scratch and the caller's spins; owns nothing
    G.d_chains = reinterpret_cast<ChainDev*>(base + vg::kScratchDesc);
    G.d_block_chain = reinterpret_cast<uint32_t*>(base + vg::kScratchMap);
    G.partials = base + part_off;
    G.arena = spins;
    G.tile = tile;
    G.total_blocks = d.nblocks;
    G.uniform_nb = d.nblocks;
    G.grid = (uint32_t)round_up(d.nblocks, vg::kNumXcd);
    G.prio = all_resident(g, tile, G.grid) ? 1u : 0u;
    G.tie_cap = 128;
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (uint64_t done = 0; done < sweeps;) {
        const uint32_t nsw = (uint32_t)std::min<uint64_t>(chunk, sweeps - done);
        d.sweep0 = sweep0 + done;
        hipLaunchKernelGGL(vg::noalloc_setup, dim3((d.nblocks + 255) / 256), dim3(256), 0, st, d,
                           G.d_chains, G.d_block_chain, d.nblocks);
        for (uint32_t s = 0; s < nsw; ++s)
            for (uint32_t c = 0; c < d.ncol; ++c) launch_zmarch_pass(G, g, st, c, s);
        hipLaunchKernelGGL(vg::ising_reduce, dim3(1, nsw), dim3(vg::kBlock), 0, st, G.d_chains,
                           (const int32_t*)G.partials, G.total_blocks, d.ncol, nsw,
                           sums + 5 * done);
        VG_HIP(hipGetLastError());
        done += nsw;
    }
    return VG_OK;
}
