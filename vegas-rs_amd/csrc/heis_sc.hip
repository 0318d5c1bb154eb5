// fp32 Heisenberg checkerboard colour pass on the simple-cubic lattice, for
// any number of chains packed in one arena.
//
// Replaces MetropolisIntegrator::step (reference src/integrator.rs:66-92:
// fresh uniform-on-sphere proposal, Spin::rand -> marsaglia, util.rs:21-34)
// and MetropolisFlipIntegrator::step (109-138: proposal -s) with the local
// energy of Exchange + UniaxialAnisotropy + Zeeman + Gauge
// (src/energy.rs:64-214): one sweep updates every colour-0 site, then every
// colour-1 site.
//
// Work item: 4 consecutive positions of one row (one aligned 16-byte load per
// component per row touched).  Random proposals draw from each site's own
// Philox calls (counter x = i), so the four sites are independent; flip
// proposals take word i & 3 of one call per item (counter x = i / 4).  The
// item runs in three phases -- fields and first proposal calls, the wave's
// queue of sites whose first call missed, then the updates -- so the rare
// further calls cost one wave-wide call per round of 64 such sites instead
// of one per site.
// All arithmetic is single IEEE fp32 operations in a fixed order (no
// contraction), so the CPU oracle reproduces every bit.
#include "heis_common.hpp"

namespace vg {

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void heis_sc_pass(const ChainDev* __restrict__ chains,
                                                    const uint32_t* __restrict__ block_chain,
                                                    float* __restrict__ arena,
                                                    int64_t* __restrict__ partials,
                                                    uint32_t colour, uint32_t s,
                                                    uint32_t total_blocks, uint32_t grid) {
    __shared__ HeisQueue hq[kBlock / 64];
    const uint32_t b = xcd_block(blockIdx.x, grid);
    if (b >= total_blocks) return;  // uniform over the workgroup
    const ChainDev& C = chains[block_chain[b]];
    if (s >= C.limit) return;  // uniform: the workgroup belongs to one chain
    const uint32_t local = (b - C.block0) * kBlock + threadIdx.x;
    const bool active = local < C.items;
    const uint64_t t = C.sweep0 + s;
    int32_t s32[4] = {0, 0, 0, 0};     // mx, my, mz, axis2 of the lane's sites
    int64_t bsum = 0;                  // bond of the lane's sites

    // Phase 1: neighbour fields, old spins and proposals of the item's sites.
    uint32_t row, i0, nvalid, pend = 0;
    // Component planes through buffer resources: 32-bit byte offsets (a
    // component plane is below 4 GiB, checked on the host), and kNoRow reads
    // 0 for an absent neighbour row.
    __amdgpu_buffer_rsrc_t rown[3], roth[3];
    {
        const uint64_t plane = C.plane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rown[c] = plane_rsrc(reinterpret_cast<const uint8_t*>(arena + C.off[colour] + c * plane),
                                 (uint32_t)(plane * 4u));
            roth[c] = plane_rsrc(
                reinterpret_cast<const uint8_t*>(arena + C.off[colour ^ 1u] + c * plane),
                (uint32_t)(plane * 4u));
        }
    }
    // Every lane runs the same code: a lane past the chain's items (the
    // tail of its last block) has no valid site and its own row offset is
    // kNoRow, so it reads zeros and its stores are dropped; a site j >=
    // nvalid of a short last segment is padding, which every kernel keeps at
    // zero.  Neither is ever accepted, so both add zero to the sums.
    row = fast_div(local, C.segs_m, C.segs_s);
    const uint32_t seg = local - row * C.segs;
    const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
    const uint32_t z = fast_div(row, C.ly_m, C.ly_s), y = row - z * ly;
    const uint32_t q = (colour + y + z) & 1u;
    i0 = seg * kHeisSeg;
    nvalid = active ? min(kHeisSeg, w - i0) : 0u;
    const uint32_t own_off = active ? (row * pitch + i0) * 4u : kNoRow;
    const uint32_t orow = row * pitch * 4u;
    float os[4][3], hs[4][3], ps[4][3], us[4];
    {
        const bool ym_ok = (y > 0) || (pbc & 2u);
        const bool yp_ok = (y + 1 < ly) || (pbc & 2u);
        const bool zm_ok = (z > 0) || (pbc & 4u);
        const bool zp_ok = (z + 1 < lz) || (pbc & 4u);
        const uint32_t r_ym = y > 0 ? row - 1 : row + (ly - 1);
        const uint32_t r_yp = y + 1 < ly ? row + 1 : row - (ly - 1);
        const uint32_t r_zm = z > 0 ? row - ly : row + (lz - 1) * ly;
        const uint32_t r_zp = z + 1 < lz ? row + ly : row - (lz - 1) * ly;

        float4 so[3], mid[3], ym[3], yp[3], zm[3], zp[3];
        float xe[3];  // the extra x neighbour (position i0-1 or i0+4)
        bool xe_ok = true;
        uint32_t xe_pos;
        if (q == 0) {
            xe_pos = i0 > 0 ? i0 - 1 : w - 1;
            xe_ok = i0 > 0 || (pbc & 1u);
        } else {
            xe_pos = i0 + kHeisSeg < w ? i0 + kHeisSeg : 0;
            xe_ok = i0 + kHeisSeg < w || (pbc & 1u);
        }
        // With q = 1 the last segment's extra neighbour is position 0: the
        // +x neighbour of position w-1 across the periodic end (0 when x is
        // open), whichever j that position has.
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            so[c] = bload4f(rown[c], own_off);
            mid[c] = bload4f(roth[c], orow + i0 * 4u);
            ym[c] = bload4f(roth[c], ym_ok ? (r_ym * pitch + i0) * 4u : kNoRow);
            yp[c] = bload4f(roth[c], yp_ok ? (r_yp * pitch + i0) * 4u : kNoRow);
            zm[c] = bload4f(roth[c], zm_ok ? (r_zm * pitch + i0) * 4u : kNoRow);
            zp[c] = bload4f(roth[c], zp_ok ? (r_zp * pitch + i0) * 4u : kNoRow);
            xe[c] = bload1f(roth[c], xe_ok ? orow + xe_pos * 4u : kNoRow);
        }

        U4 fw = U4{0u, 0u, 0u, 0u};
        if (!C.proposal) fw = flip_words(C, i0, row, t, colour);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
#pragma unroll
            for (int c = 0; c < 3; ++c) os[j][c] = f4get(so[c], j);
            const uint32_t pos = i0 + j;
            // x neighbours, -x then +x: positions pos-1, pos (q = 0) or
            // pos, pos+1 (q = 1) of the other plane's row.
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float here = f4get(mid[c], j);
                float nmx, npx;
                if (q == 0) {
                    nmx = j == 0 ? xe[c] : f4get(mid[c], j - 1);
                    npx = here;
                } else {
                    nmx = here;
                    npx = (pos + 1 == w || j == 3) ? xe[c] : f4get(mid[c], j + 1);
                }
                float v = fa(nmx, npx);
                v = fa(v, f4get(ym[c], j));
                v = fa(v, f4get(yp[c], j));
                v = fa(v, f4get(zm[c], j));
                v = fa(v, f4get(zp[c], j));
                hs[j][c] = v;
            }
            heis_propose(C, colour, pos, row, t, fw, j, os[j], us[j], ps[j], pend);
        }
        pend &= (1u << nvalid) - 1u;  // no further calls for sites that do not exist
    }

    // Phase 2: the wave's sites whose first proposal call missed.
    if (C.proposal) heis_resolve(C, hq[threadIdx.x >> 6], pend, i0, row, t, colour, ps);

    // Phase 3: accept or reject, store, sums.
    {
        float ns[3][4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            float n[3];
            heis_update(C, colour == 1, os[j], hs[j], hs[j], us[j], ps[j], n, s32, bsum,
                        j < nvalid);
            ns[0][j] = n[0];
            ns[1][j] = n[1];
            ns[2][j] = n[2];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u32x4 v;
            v.x = __float_as_uint(ns[c][0]);
            v.y = __float_as_uint(ns[c][1]);
            v.z = __float_as_uint(ns[c][2]);
            v.w = __float_as_uint(ns[c][3]);
            __builtin_amdgcn_raw_buffer_store_b128(v, rown[c], own_off, 0, 0);
        }
    }

    heis_partials_sites(s32, bsum, partials, colour, 2u, s, total_blocks, b);
}

// record = {bond, mx, my, mz, axis2}, summed over the ncol colour passes (a
// pass adds the bonds to lower colours only, so each bond counts once).
// grid = (chains, sweeps).
__global__ __launch_bounds__(256) void heis_reduce(const ChainDev* __restrict__ chains,
                                                   const int64_t* __restrict__ partials,
                                                   uint32_t total_blocks, uint32_t ncol,
                                                   uint32_t nsw, int64_t* __restrict__ records) {
    __shared__ int64_t red[5][kBlock / 64];
    const ChainDev& C = chains[blockIdx.x];
    const uint32_t s = blockIdx.y;
    if (s >= C.limit) return;  // this chain ran fewer sweeps in the block
    int64_t acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t k = threadIdx.x; k < C.nblocks; k += kBlock) {
        const uint32_t blk = C.block0 + k;
        for (uint32_t c = 0; c < ncol; ++c) {
            const int64_t* p = partials + ((uint64_t)(s * ncol + c) * total_blocks + blk) * 5u;
#pragma unroll
            for (int q = 0; q < 5; ++q) acc[q] += p[q];
        }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) red[k][wave] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        int64_t v = 0;
        for (uint32_t k = 0; k < kBlock / 64; ++k) v += red[threadIdx.x][k];
        records[((uint64_t)C.slot * nsw + s) * 5u + threadIdx.x] = v;
    }
}

}  // namespace vg
