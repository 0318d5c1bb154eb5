// Wolff single-cluster steps for int8 Ising chains on sc, bcc and fcc
// lattices (reference WolffIntegrator::step, src/integrator.rs:146-230).
//
// One step picks a seed site, grows the cluster over bonds between equal
// spins that are active with probability p = 1 - exp(-2J/T), and flips it;
// the field plays no part, as in the reference.  Randomness is keyed by the
// step and the bond, not drawn in visiting order: the seed site is
// floor(r64 * N / 2^64) for r64 = words 0,1 of Philox(0, 0, step, w(step,
// seed)), and a bond is active iff word 0 of Philox(lo, slot, step, w(step,
// bond)) is below ceil(p 2^32), (lo, slot) being its endpoint with the lower
// site id and that endpoint's neighbour slot.  The cluster is then the seed's
// connected component over active bonds, whatever order a parallel search
// visits it in, so the result is deterministic and the CPU oracle
// reproduces it.
//
// Layout (natural site order): spins int8 +-1 at off[0], a uint32 visit
// stamp per site at off[1], the cluster queue (uint32 site ids) at off[2].
// One workgroup runs all steps of one chain's sweep block: it clears the
// stamps, sums the bonds and the magnetization once, then per step searches
// level by level (the queue holds the levels back to back, a workgroup
// barrier between levels), flips the cluster, and updates the sums by the
// cluster's size and the bonds across its boundary.
#include "vg_device.hpp"

namespace vg {

namespace {

constexpr uint32_t kWolffBlock = 256;

// Coordinate v + d (d in {-1, 0, +1}) along an axis of length n; -1 when it
// leaves an open axis.
__device__ __forceinline__ int64_t wolff_axis(uint32_t v, int d, uint32_t n, bool periodic) {
    const int64_t b = (int64_t)v + d;
    if (b >= 0 && b < (int64_t)n) return b;
    if (!periodic) return -1;
    return b < 0 ? (int64_t)n - 1 : 0;
}

// fcc sublattice half-cell offsets along x, y, z as 3-bit masks.
__device__ __forceinline__ uint32_t fcc_bits(uint32_t s) {
    return s == 0 ? 0u : (s == 1 ? 6u : (s == 2 ? 5u : 3u));
}

// A site's cell coordinates and sublattice (sc: the site is its cell),
// computed once per site: the divisions by lx and ly are magic-number
// multiplications.
struct WolffSite {
    uint32_t c[3];
    uint32_t a;
};

__device__ __forceinline__ WolffSite wolff_site(const ChainDev& C, uint32_t i) {
    const uint32_t sh = C.kind == 0 ? 0u : (C.kind == 1 ? 1u : 2u);
    const uint32_t cell = i >> sh;
    const uint32_t yz = fast_div(cell, C.lx_m, C.lx_s);
    const uint32_t zz = fast_div(yz, C.ly_m, C.ly_s);
    return WolffSite{{cell - yz * C.lx, yz - zz * C.ly, zz}, i & ((1u << sh) - 1u)};
}

// Neighbour through slot k of the site at p, or -1: sc -x,+x,-y,+y,-z,+z;
// bcc slots (dz, dy, dx), dx fastest, towards cells c -+ d; fcc 4 slots per
// other sublattice b (ascending), (d1, d2) over the two axes where the
// sublattices differ, d1 fastest.  The order of the oracle's neighbour lists.
__device__ int64_t wolff_nb(const ChainDev& C, const WolffSite& p, uint32_t k) {
    const bool per[3] = {(C.pbc & 1u) != 0, (C.pbc & 2u) != 0, (C.pbc & 4u) != 0};
    const uint32_t dims[3] = {C.lx, C.ly, C.lz};
    const uint32_t* c = p.c;
    int64_t q[3];
    if (C.kind == 0) {
        const uint32_t ax = k >> 1;
        const int d = (k & 1u) ? 1 : -1;
#pragma unroll
        for (uint32_t e = 0; e < 3; ++e) {  // constant indices: no private array
            q[e] = c[e];
            if (e == ax) q[e] = wolff_axis(c[e], d, dims[e], per[e]);
        }
        if (q[0] < 0 || q[1] < 0 || q[2] < 0) return -1;
        return q[0] + (int64_t)C.lx * (q[1] + (int64_t)C.ly * q[2]);
    }
    const uint32_t sh = C.kind == 1 ? 1u : 2u;
    const uint32_t a = p.a;
    uint32_t b;
    if (C.kind == 1) {
        const int sg = a == 0 ? -1 : 1;
        const int d[3] = {(int)(k & 1u), (int)((k >> 1) & 1u), (int)((k >> 2) & 1u)};
        for (int e = 0; e < 3; ++e) {
            q[e] = wolff_axis(c[e], sg * d[e], dims[e], per[e]);
            if (q[e] < 0) return -1;
        }
        b = 1u - a;
    } else {
        const uint32_t r = k >> 2;
        b = r + (r >= a ? 1u : 0u);
        const uint32_t ba = fcc_bits(a), bb = fcc_bits(b), diff = ba ^ bb;
        const uint32_t e1 = (diff & 1u) ? 0u : 1u;          // lower differing axis
        const uint32_t e2 = (diff & 4u) ? 2u : 1u;          // higher differing axis
        const int d1 = (int)(k & 1u), d2 = (int)((k >> 1) & 1u);
        for (uint32_t e = 0; e < 3; ++e) {
            int d = (e == e1 ? d1 : 0) + (e == e2 ? d2 : 0);
            if (((bb >> e) & 1u) > ((ba >> e) & 1u)) d = -d;
            q[e] = wolff_axis(c[e], d, dims[e], per[e]);
            if (q[e] < 0) return -1;
        }
    }
    return ((q[0] + (int64_t)C.lx * (q[1] + (int64_t)C.ly * q[2])) << sh) + b;
}

// The slot of the neighbour's list that holds the same bond.
__device__ __forceinline__ uint32_t wolff_back(const ChainDev& C, uint32_t i, uint32_t k) {
    if (C.kind == 0) return k ^ 1u;
    if (C.kind == 1) return k;
    const uint32_t a = i & 3u, r = k >> 2, b = r + (r >= a ? 1u : 0u);
    return 4u * (a - (a > b ? 1u : 0u)) + (k & 3u);
}

__device__ __forceinline__ void wolff_sum2(int64_t& x, int64_t& y, int64_t (*red)[kWolffBlock / 64]) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o, 64);
        y += __shfl_xor(y, o, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        red[0][wave] = x;
        red[1][wave] = y;
    }
    __syncthreads();
    x = 0;
    y = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWolffBlock / 64; ++w) {
        x += red[0][w];
        y += red[1][w];
    }
}

}  // namespace

// grid = (chains): block b runs min(nsw, limit) steps of chain b and writes
// its records {bond, 0, 0, mag, 0} per step; chains of at most
// kWolffLdsSites sites are left to wolff_run_lds.
__global__ __launch_bounds__(kWolffBlock) void wolff_run(const ChainDev* __restrict__ chains,
                                                         uint8_t* __restrict__ arena, uint32_t nsw,
                                                         int64_t* __restrict__ records) {
    __shared__ uint32_t s_next, s_src;
    __shared__ int64_t red[2][kWolffBlock / 64];
    const ChainDev& C = chains[blockIdx.x];
    const uint32_t steps = min(nsw, C.limit);
    // uniform over the workgroup; chains of at most kWolffLdsSites sites run
    // in wolff_run_lds
    if (steps == 0 || C.nsites <= kWolffLdsSites) return;
    int8_t* const spin = reinterpret_cast<int8_t*>(arena + C.off[0]);
    uint32_t* const stamp = reinterpret_cast<uint32_t*>(arena + C.off[1]);
    uint32_t* const queue = reinterpret_cast<uint32_t*>(arena + C.off[2]);
    const uint32_t n = (uint32_t)C.nsites, z = C.z, tid = threadIdx.x;
    const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);

    // Clear the stamps; sum bonds (each once, from its lower site) and spins.
    int64_t bond = 0, mag = 0;
    for (uint32_t i = tid; i < n; i += kWolffBlock) {
        stamp[i] = 0u;
        const int32_t si = spin[i];
        mag += si;
        const WolffSite p = wolff_site(C, i);
        for (uint32_t k = 0; k < z; ++k) {
            const int64_t j = wolff_nb(C, p, k);
            if (j > (int64_t)i) bond += si * (int32_t)spin[j];
        }
    }
    wolff_sum2(bond, mag, red);

    for (uint32_t s = 0; s < steps; ++s) {
        const uint64_t t = C.sweep0 + s;
        const uint32_t sv = s + 1u;  // this step's stamp
        __syncthreads();  // the previous step's flips and queue reads are done
        if (tid == 0) {
            const U4 r = philox(U4{0u, 0u, (uint32_t)t, ctr_w(t, kWolffSeed, 0u)}, k0, k1);
            const uint64_t r64 = (uint64_t)r.x | ((uint64_t)r.y << 32);
            const uint32_t src = (uint32_t)__umul64hi(r64, (uint64_t)n);
            s_src = src;
            stamp[src] = sv;
            queue[0] = src;
        }
        __syncthreads();
        const int8_t ref = spin[s_src];
        uint32_t head = 0, tail = 1;
        while (head < tail) {  // one level per pass; tail <= n (one claim per site)
            if (tid == 0) s_next = 0u;
            __syncthreads();
            for (uint32_t q = head + tid; q < tail; q += kWolffBlock) {
                const uint32_t i = queue[q];
                const WolffSite p = wolff_site(C, i);
                for (uint32_t k = 0; k < z; ++k) {
                    const int64_t jj = wolff_nb(C, p, k);
                    if (jj < 0 || jj == (int64_t)i) continue;
                    const uint32_t j = (uint32_t)jj;
                    // a stale stamp here only costs a draw: the exchange decides
                    if (spin[j] != ref || stamp[j] == sv) continue;
                    const bool low = i < j;
                    const uint32_t lo = low ? i : j, ls = low ? k : wolff_back(C, i, k);
                    const U4 d = philox(U4{lo, ls, (uint32_t)t, ctr_w(t, kWolffBond, 0u)}, k0, k1);
                    if ((C.wolff_all || d.x < C.wolff_thr) && atomicExch(&stamp[j], sv) != sv)
                        queue[tail + atomicAdd(&s_next, 1u)] = j;
                }
            }
            __syncthreads();
            head = tail;
            tail += s_next;
            __syncthreads();
        }
        // Flip; sum the spins across the boundary (stamps read past L1).
        int64_t across = 0, dummy = 0;
        for (uint32_t q = tid; q < tail; q += kWolffBlock) {
            const uint32_t i = queue[q];
            spin[i] = (int8_t)-ref;
            const WolffSite p = wolff_site(C, i);
            for (uint32_t k = 0; k < z; ++k) {
                const int64_t jj = wolff_nb(C, p, k);
                if (jj < 0 || jj == (int64_t)i) continue;
                const uint32_t j = (uint32_t)jj;
                if (__hip_atomic_load(&stamp[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != sv)
                    across += spin[j];
            }
        }
        wolff_sum2(across, dummy, red);
        bond += -2 * (int64_t)ref * across;
        mag += -2 * (int64_t)ref * (int64_t)tail;
        if (tid == 0) {
            int64_t* r = records + ((uint64_t)C.slot * nsw + s) * 5u;
            r[0] = bond;
            r[1] = 0;
            r[2] = 0;
            r[3] = mag;
            r[4] = 0;
        }
    }
}

// The same steps with the chain's spins and visit marks in LDS, for chains
// of at most kWolffLdsSites sites: one bit per site in each array (spin bit
// set = Up), so a level's neighbour tests are LDS reads instead of global
// ones.  The cluster queue's first kWolffLdsQueue entries are in LDS too,
// the rest in global memory.  A site is claimed by
// setting its visit bit with an LDS atomic OR (the old word tells whether
// this thread set it); after the flip the cluster's visit bits are cleared
// again from the queue.  Spins are packed from the int8 array at the start
// of the sweep block and written back at its end.  The draws and the
// cluster are those of wolff_run, so both give the same steps.
// grid = (chains); blocks whose chain is larger than kWolffLdsSites return
// (wolff_run takes those).  Dynamic LDS: 2 * ceil(n / 32) + kWolffLdsQueue
// words for the largest chain n of the launch.
__global__ __launch_bounds__(kWolffBlock) void wolff_run_lds(const ChainDev* __restrict__ chains,
                                                             uint8_t* __restrict__ arena,
                                                             uint32_t nsw,
                                                             int64_t* __restrict__ records) {
    extern __shared__ uint32_t lds_bits[];
    __shared__ uint32_t s_next, s_src;
    __shared__ int64_t red[2][kWolffBlock / 64];
    const ChainDev& C = chains[blockIdx.x];
    const uint32_t steps = min(nsw, C.limit);
    if (steps == 0 || C.nsites > kWolffLdsSites) return;  // uniform over the workgroup
    int8_t* const spin = reinterpret_cast<int8_t*>(arena + C.off[0]);
    uint32_t* const queue = reinterpret_cast<uint32_t*>(arena + C.off[2]);
    const uint32_t n = (uint32_t)C.nsites, z = C.z, tid = threadIdx.x;
    const uint32_t nw = (n + 31u) >> 5;
    uint32_t* const up = lds_bits;            // spin bits
    uint32_t* const vis = lds_bits + nw;      // visit bits
    uint32_t* const lq = lds_bits + 2u * nw;  // queue entries below kWolffLdsQueue
    const auto qget = [&](uint32_t q) { return q < kWolffLdsQueue ? lq[q] : queue[q]; };
    const auto qput = [&](uint32_t q, uint32_t v) {
        if (q < kWolffLdsQueue)
            lq[q] = v;
        else
            queue[q] = v;
    };
    const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);

    // Pack the spins (each wave 64 sites per round, one byte per lane, a
    // ballot for the two words) and clear the visit bits; sum bonds (each
    // once, from its lower site) and spins from the int8 array.
    const uint32_t lane = tid & 63u;
    for (uint32_t base = (tid >> 6) * 64u; base < nw * 32u; base += kWolffBlock) {
        const uint32_t i = base + lane;
        const uint64_t m = __ballot(i < n && spin[i] > 0);
        if (lane == 0) {
            up[base >> 5] = (uint32_t)m;
            vis[base >> 5] = 0u;
            if ((base >> 5) + 1u < nw) {
                up[(base >> 5) + 1u] = (uint32_t)(m >> 32);
                vis[(base >> 5) + 1u] = 0u;
            }
        }
    }
    int64_t bond = 0, mag = 0;
    for (uint32_t i = tid; i < n; i += kWolffBlock) {
        const int32_t si = spin[i];
        mag += si;
        const WolffSite p = wolff_site(C, i);
        for (uint32_t k = 0; k < z; ++k) {
            const int64_t j = wolff_nb(C, p, k);
            if (j > (int64_t)i) bond += si * (int32_t)spin[j];
        }
    }
    wolff_sum2(bond, mag, red);  // its barriers also order the packing above

    const auto bit = [](const uint32_t* a, uint32_t i) { return (a[i >> 5] >> (i & 31u)) & 1u; };
    for (uint32_t s = 0; s < steps; ++s) {
        const uint64_t t = C.sweep0 + s;
        __syncthreads();  // the previous step's flips, clears and queue reads are done
        if (tid == 0) {
            const U4 r = philox(U4{0u, 0u, (uint32_t)t, ctr_w(t, kWolffSeed, 0u)}, k0, k1);
            const uint64_t r64 = (uint64_t)r.x | ((uint64_t)r.y << 32);
            const uint32_t src = (uint32_t)__umul64hi(r64, (uint64_t)n);
            s_src = src;
            vis[src >> 5] |= 1u << (src & 31u);
            qput(0u, src);
        }
        __syncthreads();
        const uint32_t ref = bit(up, s_src);  // the cluster's spin bit
        uint32_t head = 0, tail = 1;
        while (head < tail) {  // one level per pass; tail <= n (one claim per site)
            if (tid == 0) s_next = 0u;
            __syncthreads();
            for (uint32_t q = head + tid; q < tail; q += kWolffBlock) {
                const uint32_t i = qget(q);
                const WolffSite p = wolff_site(C, i);
                for (uint32_t k = 0; k < z; ++k) {
                    const int64_t jj = wolff_nb(C, p, k);
                    if (jj < 0 || jj == (int64_t)i) continue;
                    const uint32_t j = (uint32_t)jj;
                    if (bit(up, j) != ref || bit(vis, j)) continue;
                    const bool low = i < j;
                    const uint32_t lo = low ? i : j, ls = low ? k : wolff_back(C, i, k);
                    const U4 d = philox(U4{lo, ls, (uint32_t)t, ctr_w(t, kWolffBond, 0u)}, k0, k1);
                    if (C.wolff_all || d.x < C.wolff_thr) {
                        const uint32_t mb = 1u << (j & 31u);
                        if ((atomicOr(&vis[j >> 5], mb) & mb) == 0u)
                            qput(tail + atomicAdd(&s_next, 1u), j);
                    }
                }
            }
            __syncthreads();
            head = tail;
            tail += s_next;
            __syncthreads();
        }
        // Flip; sum the spins across the boundary (neighbours outside the
        // cluster keep their spins while the cluster flips).
        int64_t across = 0, dummy = 0;
        for (uint32_t q = tid; q < tail; q += kWolffBlock) {
            const uint32_t i = qget(q);
            atomicXor(&up[i >> 5], 1u << (i & 31u));
            const WolffSite p = wolff_site(C, i);
            for (uint32_t k = 0; k < z; ++k) {
                const int64_t jj = wolff_nb(C, p, k);
                if (jj < 0 || jj == (int64_t)i) continue;
                const uint32_t j = (uint32_t)jj;
                if (!bit(vis, j)) across += bit(up, j) ? 1 : -1;
            }
        }
        wolff_sum2(across, dummy, red);  // every visit-bit read is done after it
        for (uint32_t q = tid; q < tail; q += kWolffBlock) {
            const uint32_t i = qget(q);
            atomicAnd(&vis[i >> 5], ~(1u << (i & 31u)));
        }
        const int64_t sref = ref ? 1 : -1;
        bond += -2 * sref * across;
        mag += -2 * sref * (int64_t)tail;
        if (tid == 0) {
            int64_t* r = records + ((uint64_t)C.slot * nsw + s) * 5u;
            r[0] = bond;
            r[1] = 0;
            r[2] = 0;
            r[3] = mag;
            r[4] = 0;
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kWolffBlock) spin[i] = bit(up, i) ? (int8_t)1 : (int8_t)-1;
}

// Random initial state in natural order, the same draws as the colour-split
// kernels: site n at (colour c, row, position i) is Up iff word (i & 3) of
// Philox(i >> 2, row, 0, w(0, init, c)) is below 2^31.
__global__ void wolff_init(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= C.nsites) return;
    uint32_t c, row, i;
    if (C.kind == 0) {
        const uint32_t x = (uint32_t)(n % C.lx), y = (uint32_t)((n / C.lx) % C.ly);
        const uint32_t zz = (uint32_t)(n / ((uint64_t)C.lx * C.ly));
        c = (x + y + zz) & 1u;
        row = y + C.ly * zz;
        i = x >> 1;
    } else {
        const uint32_t sh = C.kind == 1 ? 1u : 2u;
        const uint64_t cell = n >> sh;
        const uint32_t x = (uint32_t)(cell % C.lx), y = (uint32_t)((cell / C.lx) % C.ly);
        const uint32_t zz = (uint32_t)(cell / ((uint64_t)C.lx * C.ly));
        c = (uint32_t)(n & ((1u << sh) - 1u));
        row = y + C.ly * zz;
        i = x;
    }
    const U4 r = philox(U4{i >> 2, row, 0u, ctr_w(0, kIsingInit, c)}, (uint32_t)C.seed,
                        (uint32_t)(C.seed >> 32));
    reinterpret_cast<int8_t*>(arena + C.off[0])[n] = (u4_word(r, i & 3u) >> 31) == 0 ? 1 : -1;
}

}  // namespace vg
