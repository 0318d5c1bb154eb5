// Int8 Ising checkerboard colour pass on the simple-cubic lattice, for any
// number of chains packed in one arena (a single chain is a batch of one).
//
// Replaces the inner loop of MetropolisFlipIntegrator::step and
// MetropolisIntegrator::step (reference src/integrator.rs:66-92, 109-138) with
// the Exchange + Zeeman local energy of src/energy.rs:147-151, 194-206:
// instead of N attempts at uniformly random sites, one sweep updates every
// colour-0 site, then every colour-1 site.  A site's flip is accepted when
// dE <= 0, or else when its 32-bit draw is below floor-free threshold
// ceil(exp(-dE/T) * 2^32) -- the reference's `u < exp(-dE/T)` rule
// (integrator.rs:127-134) with u quantised to 2^-32.
//
// Work item: 16 consecutive positions of one row of the colour plane (one
// aligned 16-byte load per row touched).  Neighbour counts are summed four
// sites per 32-bit register.  The acceptance draw of each site is split: its
// high byte comes from one Philox call shared by the item's 16 sites; only
// when that byte equals the high byte of the threshold (probability 2^-8)
// is the low 24 bits' Philox call made.  The result equals comparing the full
// 32-bit draw.
#include "vg_device.hpp"

namespace vg {

__device__ __forceinline__ uint32_t sel4(uint32_t k, uint32_t a, uint32_t b, uint32_t c,
                                         uint32_t d) {
    return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d));
}

__device__ __forceinline__ uint32_t set_byte(uint32_t word, uint32_t byte, uint32_t value) {
    const uint32_t sh = 8u * byte;
    return (word & ~(0xFFu << sh)) | (value << sh);
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Loads one neighbour row segment of the other colour plane (or marks it
// absent on an open boundary).
__device__ __forceinline__ uint4 load_seg(const uint8_t* plane, int64_t row, uint32_t pitch,
                                          uint32_t i0, bool present) {
    if (!present) return make_uint4(0, 0, 0, 0);
    return *reinterpret_cast<const uint4*>(plane + row * (int64_t)pitch + i0);
}

__global__ __launch_bounds__(256) void ising_sc_pass(const ChainDev* __restrict__ chains,
                                                     const uint32_t* __restrict__ block_chain,
                                                     uint8_t* __restrict__ arena,
                                                     int32_t* __restrict__ partials,
                                                     uint32_t colour, uint32_t s,
                                                     uint32_t total_blocks, uint32_t grid) {
    __shared__ uint32_t lut[34];
    __shared__ int32_t red[2][kBlock / 64];

    const uint32_t b = xcd_block(blockIdx.x, grid);
    if (b >= total_blocks) return;  // uniform over the workgroup
    const ChainDev& C = chains[block_chain[b]];
    if (threadIdx.x < 34) lut[threadIdx.x] = (&C.thr[0][0])[threadIdx.x];
    __syncthreads();

    const uint32_t local = (b - C.block0) * kBlock + threadIdx.x;
    int32_t m_part = 0, b_part = 0;

    if (local < C.items) {
        const uint32_t row = local / C.segs;
        const uint32_t seg = local - row * C.segs;
        const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
        const uint32_t y = row % ly, z = row / ly;
        const uint32_t q = (colour + y + z) & 1u;  // own sites sit at x = 2i + q
        const uint32_t i0 = seg * kIsingSeg;
        const uint32_t nvalid = min(kIsingSeg, w - i0);

        uint8_t* own_plane = arena + C.off[colour];
        const uint8_t* oth = arena + C.off[colour ^ 1u];
        uint4* own_ptr = reinterpret_cast<uint4*>(own_plane + (uint64_t)row * pitch + i0);
        const uint4 o4 = *own_ptr;
        uint4 m4 = load_seg(oth, row, pitch, i0, true);

        // Neighbour rows along y and z (same positions i in the other plane).
        uint32_t gu = 0;  // absent neighbour slots common to all 16 sites
        const bool ym_ok = (y > 0) || (pbc & 2u);
        const bool yp_ok = (y + 1 < ly) || (pbc & 2u);
        const bool zm_ok = (z > 0) || (pbc & 4u);
        const bool zp_ok = (z + 1 < lz) || (pbc & 4u);
        const int64_t r_ym = y > 0 ? (int64_t)row - 1 : (int64_t)row + (ly - 1);
        const int64_t r_yp = y + 1 < ly ? (int64_t)row + 1 : (int64_t)row - (ly - 1);
        const int64_t r_zm = z > 0 ? (int64_t)row - ly : (int64_t)row + (int64_t)(lz - 1) * ly;
        const int64_t r_zp =
            z + 1 < lz ? (int64_t)row + ly : (int64_t)row - (int64_t)(lz - 1) * ly;
        const uint4 ym4 = load_seg(oth, r_ym, pitch, i0, ym_ok);
        const uint4 yp4 = load_seg(oth, r_yp, pitch, i0, yp_ok);
        const uint4 zm4 = load_seg(oth, r_zm, pitch, i0, zm_ok);
        const uint4 zp4 = load_seg(oth, r_zp, pitch, i0, zp_ok);
        gu = (ym_ok ? 0u : 1u) + (yp_ok ? 0u : 1u) + (zm_ok ? 0u : 1u) + (zp_ok ? 0u : 1u);

        // Along x the two neighbours of position i are i-1,i (q = 0) or i,i+1
        // (q = 1) in the other plane's same row.
        const uint8_t* orow = oth + (uint64_t)row * pitch;
        const bool px = pbc & 1u;
        uint32_t m[4] = {m4.x, m4.y, m4.z, m4.w};
        uint32_t xs[4];
        uint32_t gx_pos = 0xFFFFFFFFu;  // position whose x neighbour is absent
        if (q == 0) {
            uint32_t extra = 0;
            if (i0 > 0)
                extra = orow[i0 - 1];
            else if (px)
                extra = orow[w - 1];
            else
                gx_pos = 0;
            xs[0] = __builtin_amdgcn_alignbyte(m[0], extra << 24, 3);
            xs[1] = __builtin_amdgcn_alignbyte(m[1], m[0], 3);
            xs[2] = __builtin_amdgcn_alignbyte(m[2], m[1], 3);
            xs[3] = __builtin_amdgcn_alignbyte(m[3], m[2], 3);
        } else {
            uint32_t extra = 0;
            if (i0 + kIsingSeg < w) {
                extra = orow[i0 + kIsingSeg];
            } else {
                // The last position w-1 needs position w, i.e. position 0
                // across the periodic boundary, or nothing when x is open.
                const uint32_t wrap = px ? (uint32_t)orow[0] : 0u;
                const uint32_t pw = w - i0;  // 1..16
                if (!px) gx_pos = pw - 1;
                if (pw == kIsingSeg) {
                    extra = wrap;
                } else {
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k)
                        if ((pw >> 2) == k) m[k] = set_byte(m[k], pw & 3u, wrap);
                }
            }
            xs[0] = __builtin_amdgcn_alignbyte(m[1], m[0], 1);
            xs[1] = __builtin_amdgcn_alignbyte(m[2], m[1], 1);
            xs[2] = __builtin_amdgcn_alignbyte(m[3], m[2], 1);
            xs[3] = __builtin_amdgcn_alignbyte(extra, m[3], 1);
        }
        const uint32_t ymv[4] = {ym4.x, ym4.y, ym4.z, ym4.w};
        const uint32_t ypv[4] = {yp4.x, yp4.y, yp4.z, yp4.w};
        const uint32_t zmv[4] = {zm4.x, zm4.y, zm4.z, zm4.w};
        const uint32_t zpv[4] = {zp4.x, zp4.y, zp4.z, zp4.w};
        const uint32_t o[4] = {o4.x, o4.y, o4.z, o4.w};
        // S = 2 * (up neighbours) + (absent neighbour slots), per byte <= 12.
        uint32_t S[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // m[k] differs from the loaded row only at a padding position.
            const uint32_t cnt = m[k] + xs[k] + ymv[k] + ypv[k] + zmv[k] + zpv[k];
            S[k] = (cnt << 1) + gu * 0x01010101u;
            if ((gx_pos >> 2) == (uint32_t)k) S[k] += 1u << (8u * (gx_pos & 3u));
        }

        const uint64_t t = C.sweep0 + s;
        const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);
        const U4 ra = philox(U4{i0 >> 4, row, (uint32_t)t, ctr_w(t, kIsingA, colour)}, k0, k1);
        const uint32_t rw[4] = {ra.x, ra.y, ra.z, ra.w};
        uint32_t cand = 0xFFFFu;  // flip proposal: every site is a candidate
        if (C.proposal) {
            const U4 rp = philox(U4{i0 >> 7, row, (uint32_t)t, ctr_w(t, kIsingProposal, colour)},
                                 k0, k1);
            const uint32_t word = u4_word(rp, (i0 >> 5) & 3u);
            const uint32_t bits = (word >> (i0 & 31u)) & 0xFFFFu;
            // Spin::rand proposes Up when its draw is below 1/2 (state.rs:76-84):
            // bit 0 -> Up.  A proposal equal to the current spin changes nothing.
            uint32_t ob = 0;
#pragma unroll
            for (int j = 0; j < 16; ++j) ob |= ((o[j >> 2] >> (8 * (j & 3))) & 1u) << j;
            cand = ~(bits ^ ob) & 0xFFFFu;
        }
        const uint32_t always0 = C.always[0], always1 = C.always[1];
        uint32_t flip = 0, tie = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = j >> 2, sh = 8 * (j & 3);
            const uint32_t sj = (S[k] >> sh) & 0xFFu;
            const uint32_t bj = (o[k] >> sh) & 1u;
            const uint32_t alw = ((bj ? always1 : always0) >> sj) & 1u;
            const uint32_t th_hi = lut[bj * 17u + sj] >> 24;
            const uint32_t r8 = (rw[k] >> sh) & 0xFFu;
            flip |= (alw | (uint32_t)(r8 < th_hi)) << j;
            tie |= ((alw ^ 1u) & (uint32_t)(r8 == th_hi)) << j;
        }
        const uint32_t vmask = (nvalid >= 16) ? 0xFFFFu : ((1u << nvalid) - 1u);
        flip &= cand & vmask;
        tie &= cand & vmask;
        while (tie) {
            const uint32_t j = __builtin_ctz(tie);
            const uint32_t g = j >> 2;  // group of 4 positions sharing one call
            const U4 rb = philox(U4{(i0 >> 2) + g, row, (uint32_t)t, ctr_w(t, kIsingB, colour)},
                                 k0, k1);
            const uint32_t sg = sel4(g, S[0], S[1], S[2], S[3]);
            const uint32_t og = sel4(g, o[0], o[1], o[2], o[3]);
#pragma unroll
            for (uint32_t jj = 0; jj < 4; ++jj) {
                const uint32_t pos = 4 * g + jj;
                if ((tie >> pos) & 1u) {
                    const uint32_t sj = (sg >> (8 * jj)) & 0xFFu;
                    const uint32_t bj = (og >> (8 * jj)) & 1u;
                    const uint32_t th = lut[bj * 17u + sj];
                    const uint32_t lo24 = u4_word(rb, jj) >> 8;
                    flip |= (uint32_t)(lo24 < (th & 0xFFFFFFu)) << pos;
                }
            }
            tie &= ~(0xFu << (4 * g));
        }

        uint32_t n[4];
        int32_t up = 0, bond = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t fm = 0;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) fm |= ((flip >> (4 * k + jj)) & 1u) << (8 * jj);
            n[k] = o[k] ^ fm;
            up += __popc(n[k]);
        }
        if (flip) *own_ptr = make_uint4(n[0], n[1], n[2], n[3]);
        m_part = 2 * up - (int32_t)nvalid;
        if (colour == 1) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int k = j >> 2, sh = 8 * (j & 3);
                const int32_t sj = (int32_t)((S[k] >> sh) & 0xFFu) - 6;
                const int32_t nb = (int32_t)((n[k] >> sh) & 1u);
                bond += ((uint32_t)j < nvalid) ? (nb ? sj : -sj) : 0;
            }
        }
        b_part = bond;
    }

    m_part = wave_sum(m_part);
    b_part = wave_sum(b_part);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = m_part;
        red[1][wave] = b_part;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t mm = 0, bb = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBlock / 64; ++k) {
            mm += red[0][k];
            bb += red[1][k];
        }
        int32_t* p = partials + ((uint64_t)(s * 2u + colour) * total_blocks + b) * 2u;
        p[0] = mm;
        p[1] = bb;
    }
}

// Sums the pass partials of each (chain, sweep) into its record:
// record = {bond, 0, 0, mag, 0} (int64).  grid = (chains, sweeps).
__global__ __launch_bounds__(256) void ising_reduce(const ChainDev* __restrict__ chains,
                                                    const int32_t* __restrict__ partials,
                                                    uint32_t total_blocks, uint32_t nsw,
                                                    int64_t* __restrict__ records) {
    __shared__ int64_t red[2][kBlock / 64];
    const ChainDev& C = chains[blockIdx.x];
    const uint32_t s = blockIdx.y;
    int64_t mm = 0, bb = 0;
    for (uint32_t k = threadIdx.x; k < C.nblocks; k += kBlock) {
        const uint32_t blk = C.block0 + k;
        const int32_t* p0 = partials + ((uint64_t)(s * 2u + 0u) * total_blocks + blk) * 2u;
        const int32_t* p1 = partials + ((uint64_t)(s * 2u + 1u) * total_blocks + blk) * 2u;
        mm += (int64_t)p0[0] + (int64_t)p1[0];
        bb += (int64_t)p1[1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mm += __shfl_xor(mm, o, 64);
        bb += __shfl_xor(bb, o, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = mm;
        red[1][wave] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t m = 0, bo = 0;
        for (uint32_t k = 0; k < kBlock / 64; ++k) {
            m += red[0][k];
            bo += red[1][k];
        }
        int64_t* r = records + ((uint64_t)C.slot * nsw + s) * 5u;
        r[0] = bo;
        r[1] = 0;
        r[2] = 0;
        r[3] = m;
        r[4] = 0;
    }
}

}  // namespace vg
