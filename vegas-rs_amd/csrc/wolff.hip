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

// Neighbour of site i through slot k, or -1: sc -x,+x,-y,+y,-z,+z; bcc
// slots (dz, dy, dx), dx fastest, towards cells c -+ d; fcc 4 slots per
// other sublattice b (ascending), (d1, d2) over the two axes where the
// sublattices differ, d1 fastest.  The order of the oracle's neighbour lists.
__device__ int64_t wolff_nb(const ChainDev& C, uint32_t i, uint32_t k) {
    const bool per[3] = {(C.pbc & 1u) != 0, (C.pbc & 2u) != 0, (C.pbc & 4u) != 0};
    const uint32_t dims[3] = {C.lx, C.ly, C.lz};
    if (C.kind == 0) {
        const uint32_t c[3] = {i % C.lx, (i / C.lx) % C.ly, i / (C.lx * C.ly)};
        const uint32_t ax = k >> 1;
        const int d = (k & 1u) ? 1 : -1;
        int64_t q[3] = {c[0], c[1], c[2]};
        q[ax] = wolff_axis(c[ax], d, dims[ax], per[ax]);
        if (q[ax] < 0) return -1;
        return q[0] + (int64_t)C.lx * (q[1] + (int64_t)C.ly * q[2]);
    }
    const uint32_t sh = C.kind == 1 ? 1u : 2u;
    const uint32_t cell = i >> sh, a = i & ((1u << sh) - 1u);
    const uint32_t c[3] = {cell % C.lx, (cell / C.lx) % C.ly, cell / (C.lx * C.ly)};
    int64_t q[3];
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
// its records {bond, 0, 0, mag, 0} per step.
__global__ __launch_bounds__(kWolffBlock) void wolff_run(const ChainDev* __restrict__ chains,
                                                         uint8_t* __restrict__ arena, uint32_t nsw,
                                                         int64_t* __restrict__ records) {
    __shared__ uint32_t s_next, s_src;
    __shared__ int64_t red[2][kWolffBlock / 64];
    const ChainDev& C = chains[blockIdx.x];
    const uint32_t steps = min(nsw, C.limit);
    if (steps == 0) return;  // uniform over the workgroup
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
        for (uint32_t k = 0; k < z; ++k) {
            const int64_t j = wolff_nb(C, i, k);
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
                for (uint32_t k = 0; k < z; ++k) {
                    const int64_t jj = wolff_nb(C, i, k);
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
            for (uint32_t k = 0; k < z; ++k) {
                const int64_t jj = wolff_nb(C, i, k);
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
