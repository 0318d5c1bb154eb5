// Int8 Ising checkerboard colour pass on the simple-cubic lattice, for any
// number of chains packed in one arena (a single chain is a batch of one).
//
// Replaces the inner loop of MetropolisFlipIntegrator::step and
// MetropolisIntegrator::step (reference src/integrator.rs:66-92, 109-138) with
// the Exchange + Zeeman local energy of src/energy.rs:147-151, 194-206:
// instead of N attempts at uniformly random sites, one sweep updates every
// colour-0 site, then every colour-1 site.  A site's flip is accepted when
// dE <= 0, or else when its 32-bit draw k satisfies k < ceil(exp(-dE/T)*2^32)
// -- the reference's `u < exp(-dE/T)` rule (integrator.rs:127-134) with u
// quantised to 2^-32.
//
// Work item: one 16-position segment of a row, marched through K (the tile,
// a template parameter picked by the host) consecutive z planes.  Along z the row below is the previous row's
// neighbour above, so each plane costs four 16-byte loads (own, y-1, y+1,
// z+1) instead of six.  Items of one wave share the parity of y, so the
// direction of the x-neighbour byte shift (which alternates with y+z) is
// uniform across the wave.
//
// Per site everything is done four sites per 32-bit register: neighbour
// counts by byte-wise adds, the acceptance class lookup by v_perm_b32 on
// 8-entry byte tables, and the comparison of the draw's high byte with the
// threshold's high byte by borrow-free byte arithmetic.  The high bytes of a
// segment's 16 draws come from one Philox call.  When a draw's high byte
// equals the threshold's (probability 2^-8) the low 24 bits decide; those
// ties are rare per row but common per wave, so a row with ties only leaves a
// record (slot by ballot) in the wave's LDS queue; once the wave has finished
// its planes, one lane per record resolves its ties, writes the flipped bytes
// and corrects the wave's sums.  The result equals comparing the full 32-bit
// draw.  No workgroup barrier is needed after the tables are loaded.
#include "ising_swar.hpp"

namespace vg {

constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kTieRows = 64;     // tie records per wave (one per row with ties)
constexpr uint32_t kTieWords = 10;

__device__ __forceinline__ uint32_t pick4(uint32_t k, uint32_t a, uint32_t b, uint32_t c,
                                          uint32_t d) {
    return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d));
}

template <uint32_t K, bool BITS>
__global__ __launch_bounds__(256) void ising_sc_pass(const ChainDev* __restrict__ chains,
                                                     const uint32_t* __restrict__ block_chain,
                                                     uint8_t* __restrict__ arena,
                                                     int32_t* __restrict__ partials,
                                                     uint32_t colour, uint32_t s,
                                                     uint32_t total_blocks, uint32_t grid,
                                                     uint32_t opts) {
    __shared__ uint32_t lut[2 * kThrS];
    __shared__ uint32_t tab[20];  // t8, [g][b][2]
    // per wave: kTieRows records of {meta, tie mask, cnt[4], own[4]}
    __shared__ uint32_t tie_q[kWaves][kTieRows][kTieWords];
    // opts: bit 0 = falling wave priorities; bits 8-15 = records the tie
    // queue may hold (at most kTieRows; fewer only to exercise the overflow
    // path in tests -- results do not depend on it).
    const uint32_t prio = opts & 1u;
    const uint32_t qcap = min((opts >> 8) & 0xFFu, kTieRows);

    const uint32_t b = xcd_block(blockIdx.x, grid);
    if (b >= total_blocks) return;  // uniform over the workgroup
    const ChainDev& C = chains[block_chain[b]];
    if (s >= C.limit) return;  // uniform: the workgroup belongs to one chain
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid < 2 * kThrS) lut[tid] = (&C.thr[0][0])[tid];
    if (tid >= 64 && tid < 84) tab[tid - 64] = (&C.t8[0][0][0])[tid - 64];
    __syncthreads();

    const uint64_t t = C.sweep0 + s;
    const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);
    const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
    const uint32_t always0 = C.always[0], always1 = C.always[1];
    // A colour plane is below 4 GiB (checked on the host): 32-bit offsets.
    uint8_t* const own_plane = arena + C.off[colour];
    const uint8_t* const oth = arena + C.off[colour ^ 1u];
    const __amdgpu_buffer_rsrc_t rw_own = plane_rsrc(own_plane, (uint32_t)C.plane);
    const __amdgpu_buffer_rsrc_t ro = plane_rsrc(oth, (uint32_t)C.plane);
    const uint32_t zstep = ly * pitch;  // bytes between planes z and z+1
    const uint32_t local = (b - C.block0) * kBlock + tid;
    uint32_t (*const q_wave)[kTieWords] = tie_q[wave];
    uint32_t nrec = 0;  // records queued by this wave, as counted by this lane

    int32_t m_part = 0, b_part = 0;
    uint32_t seg = 0, y = 0, zc = 0;
    if (local < C.items) {
        const uint32_t r = fast_div(local, C.segs_m, C.segs_s);
        seg = local - r * C.segs;
        const uint32_t r2 = fast_div(r, C.hy_m, C.hy_s);
        y = 2u * (r - r2 * C.hy) + (r2 & 1u);
        zc = r2 >> 1;
    }
    if (local < C.items && y < ly) {
        const uint32_t z0 = zc * K;
        const uint32_t nk = min(K, lz - z0);
        const uint32_t i0 = seg * kIsingSeg;
        const uint32_t nvalid = min(kIsingSeg, w - i0);
        const uint32_t vbits = nvalid >= 16 ? 0xFFFFu : ((1u << nvalid) - 1u);
        uint32_t validH[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) validH[k] = nibble_bytes(vbits >> (4 * k)) << 7;
        const bool px = pbc & 1u, py = pbc & 2u, pz = pbc & 4u;
        const bool ym_ok = y > 0 || py;
        const bool yp_ok = y + 1 < ly || py;
        // offsets of rows y-1 and y+1 relative to row y (wrapping when periodic)
        const int32_t dym = y > 0 ? -(int32_t)pitch : (int32_t)((ly - 1) * pitch);
        const int32_t dyp = y + 1 < ly ? (int32_t)pitch : -(int32_t)((ly - 1) * pitch);
        const uint32_t gy = (ym_ok ? 0u : 1u) + (yp_ok ? 0u : 1u);
        // x neighbour outside the 16 positions: i0-1 (q = 0) or i0+16 (q = 1)
        const bool xl_ok = i0 > 0 || px;
        const uint32_t xl_pos = i0 > 0 ? i0 - 1 : w - 1;
        const uint32_t pw = w - i0;  // positions left in the row, 1..16 in the last segment
        const bool xr_in = i0 + kIsingSeg < w;
        uint32_t off = (y + ly * z0) * pitch + seg_off<BITS>(i0);  // row (y, z0), this segment
        // the row window along z: row (y, z-1) and row (y, z)
        const bool zm0_ok = z0 > 0 || pz;
        u32x4 below =
            load_seg<BITS>(ro, !zm0_ok ? kNoRow : (z0 > 0 ? off - zstep : off + (lz - 1) * zstep));
        u32x4 here = load_seg<BITS>(ro, off);
        bool below_ok = zm0_ok;

#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            if (k >= nk) break;
            // The SIMD arbiter favours older waves; when every wave of the
            // pass is resident at once, a wave's priority falls as it
            // marches, so the waves of a SIMD progress together and the pass
            // does not end in a tail of a few young waves.  (With a second
            // round of waves, staggered ends let it start early instead.)
            if (prio) set_prio(((K - 1u - k) * 4u) / K);
            const uint32_t z = z0 + k;
            const uint32_t row = y + ly * z;
            const bool zp_ok = z + 1 < lz || pz;
            const u32x4 o4 = load_seg<BITS>(rw_own, off);
            const u32x4 ym4 = load_seg<BITS>(ro, ym_ok ? off + dym : kNoRow);
            const u32x4 yp4 = load_seg<BITS>(ro, yp_ok ? off + dyp : kNoRow);
            const u32x4 above = load_seg<BITS>(
                ro, !zp_ok ? kNoRow : (z + 1 < lz ? off + zstep : off - (lz - 1) * zstep));
            const uint32_t g = gy + (below_ok ? 0u : 1u) + (zp_ok ? 0u : 1u);

            // x neighbours of position i: i-1,i (q = 0) or i,i+1 (q = 1) in
            // the other plane's same row.  The byte outside the segment is
            // position i0-1 (q = 0), or i0+16, or position 0 across the
            // periodic end (q = 1); kNoRow reads zero when x is open.
            const uint32_t q = (colour + y + z) & 1u;
            const uint32_t orow = off - seg_off<BITS>(i0);
            uint32_t m[4] = {here.x, here.y, here.z, here.w};
            uint32_t xs[4];
            uint32_t gx = 0xFFFFFFFFu;  // position whose x neighbour is absent
            if (q == 0) {
                const uint32_t extra = load_pos<BITS>(ro, orow, xl_pos, xl_ok);
                if (!xl_ok) gx = 0;
                xs[0] = __builtin_amdgcn_alignbyte(m[0], extra << 24, 3);
                xs[1] = __builtin_amdgcn_alignbyte(m[1], m[0], 3);
                xs[2] = __builtin_amdgcn_alignbyte(m[2], m[1], 3);
                xs[3] = __builtin_amdgcn_alignbyte(m[3], m[2], 3);
            } else {
                uint32_t extra = load_pos<BITS>(ro, orow, xr_in ? i0 + kIsingSeg : 0, xr_in || px);
                if (!xr_in) {
                    // Position w-1 needs position w: position 0 across the
                    // periodic boundary, or nothing when x is open.
                    if (!px) gx = pw - 1;
                    if (pw != kIsingSeg) {
#pragma unroll
                        for (uint32_t kk = 0; kk < 4; ++kk)
                            if ((pw >> 2) == kk) {
                                const uint32_t sh = 8u * (pw & 3u);
                                m[kk] = (m[kk] & ~(0xFFu << sh)) | (extra << sh);
                            }
                        extra = 0;
                    }
                }
                xs[0] = __builtin_amdgcn_alignbyte(m[1], m[0], 1);
                xs[1] = __builtin_amdgcn_alignbyte(m[2], m[1], 1);
                xs[2] = __builtin_amdgcn_alignbyte(m[3], m[2], 1);
                xs[3] = __builtin_amdgcn_alignbyte(extra, m[3], 1);
            }
            const uint32_t o[4] = {o4.x, o4.y, o4.z, o4.w};
            const uint32_t dn[4] = {below.x, below.y, below.z, below.w};
            const uint32_t up4[4] = {above.x, above.y, above.z, above.w};
            const uint32_t ymv[4] = {ym4.x, ym4.y, ym4.z, ym4.w};
            const uint32_t ypv[4] = {yp4.x, yp4.y, yp4.z, yp4.w};
            const uint32_t t0lo = tab[g * 4 + 0], t0hi = tab[g * 4 + 1];
            const uint32_t t1lo = tab[g * 4 + 2], t1hi = tab[g * 4 + 3];

            const U4 ra = philox(U4{i0 >> 4, row, (uint32_t)t, ctr_w(t, kIsingA, colour)}, k0, k1);
            const uint32_t rw[4] = {ra.x, ra.y, ra.z, ra.w};
            uint32_t live[4] = {validH[0], validH[1], validH[2], validH[3]};
            if (C.proposal) {
                const U4 rp = philox(
                    U4{i0 >> 7, row, (uint32_t)t, ctr_w(t, kIsingProposal, colour)}, k0, k1);
                const uint32_t bits = (u4_word(rp, (i0 >> 5) & 3u) >> (i0 & 31u)) & 0xFFFFu;
                // Spin::rand proposes Up when its draw is below 1/2
                // (state.rs:76-84): proposal bit 0 -> Up.  A proposal equal
                // to the current spin changes nothing.
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    live[kk] &= (~(nibble_bytes(bits >> (4 * kk)) ^ o[kk]) & kOne) << 7;
            }
            uint32_t cnt[4], accH[4], tieH[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                cnt[kk] = m[kk] + xs[kk] + ymv[kk] + ypv[kk] + dn[kk] + up4[kk];  // 0..6
                const uint32_t bm = __builtin_amdgcn_perm(0u, 0xFF00u, o[kk]);   // 0xFF where Up
                const uint32_t th0 = __builtin_amdgcn_perm(t0hi, t0lo, cnt[kk]);
                const uint32_t th1 = __builtin_amdgcn_perm(t1hi, t1lo, cnt[kk]);
                const uint32_t th = (th1 & bm) | (th0 & ~bm);
                accH[kk] = bytes_lt(rw[kk], th) & live[kk];
                tieH[kk] = bytes_eq(rw[kk], th) & live[kk];
            }
            // The one position with an absent x neighbour (open x boundary)
            // has S = 2c + g + 1, which the tables above do not cover.
            if (gx < nvalid) {
                const uint32_t kk = gx >> 2, sh = 8u * (gx & 3u);
                const uint32_t c = (pick4(kk, cnt[0], cnt[1], cnt[2], cnt[3]) >> sh) & 0xFFu;
                const uint32_t bj = (pick4(kk, o[0], o[1], o[2], o[3]) >> sh) & 1u;
                const uint32_t S = 2u * c + g + 1u;
                const uint32_t alw = ((bj ? always1 : always0) >> S) & 1u;
                const uint32_t th_hi = alw ? 0xFFu : lut[bj * kThrS + S] >> 24;
                const uint32_t r8 = (pick4(kk, rw[0], rw[1], rw[2], rw[3]) >> sh) & 0xFFu;
                const uint32_t lv = (pick4(kk, live[0], live[1], live[2], live[3]) >> sh) & 0x80u;
                const uint32_t a = (lv && r8 < th_hi) ? 0x80u : 0u;
                const uint32_t ti = (lv && r8 == th_hi) ? 0x80u : 0u;
#pragma unroll
                for (uint32_t q2 = 0; q2 < 4; ++q2)
                    if (q2 == kk) {
                        accH[q2] = (accH[q2] & ~(0x80u << sh)) | (a << sh);
                        tieH[q2] = (tieH[q2] & ~(0x80u << sh)) | (ti << sh);
                    }
            }
            uint32_t n[4], any = 0;
            int32_t upc = 0;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint32_t f = accH[kk] >> 7;
                any |= f;
                n[kk] = o[kk] ^ f;
                upc += __popc(n[kk]);
            }
            if (any) {
                if (BITS) {
                    __builtin_amdgcn_raw_buffer_store_b16(
                        (unsigned short)pack16(n[0], n[1], n[2], n[3]), rw_own, off, 0, 0);
                } else {
                    u32x4 nv;
                    nv.x = n[0];
                    nv.y = n[1];
                    nv.z = n[2];
                    nv.w = n[3];
                    __builtin_amdgcn_raw_buffer_store_b128(nv, rw_own, off, 0, 0);
                }
            }
            m_part += 2 * upc - (int32_t)nvalid;
            if (colour == 1) {
                // sum_j s_j (S_j - 6), s = 2n - 1, S = 2c + g (+1 at an open x end)
                uint32_t nc = 0, cs = 0;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const uint32_t cv = cnt[kk] & ((validH[kk] >> 7) * 0xFFu);
                    nc = __builtin_amdgcn_udot4(n[kk], cv, nc, false);
                    cs = __builtin_amdgcn_udot4(cv, kOne, cs, false);
                }
                const int32_t g6 = (int32_t)g - 6;
                b_part += 4 * (int32_t)nc + 2 * g6 * upc - 2 * (int32_t)cs - g6 * (int32_t)nvalid;
                if (gx < nvalid) {
                    const uint32_t kk = gx >> 2, sh = 8u * (gx & 3u);
                    const int32_t nj = (int32_t)((pick4(kk, n[0], n[1], n[2], n[3]) >> sh) & 1u);
                    b_part += 2 * nj - 1;
                }
            }
            // Rows with ties leave one record each in the wave's queue (their
            // tied bytes were written unchanged above); slots by ballot.
            // tie flags -> 16-bit mask (byte MSBs gathered by a multiply)
            const uint32_t tmask = ((((tieH[0] >> 7) & kOne) * 0x10204080u) >> 28) |
                                   (((((tieH[1] >> 7) & kOne) * 0x10204080u) >> 28) << 4) |
                                   (((((tieH[2] >> 7) & kOne) * 0x10204080u) >> 28) << 8) |
                                   (((((tieH[3] >> 7) & kOne) * 0x10204080u) >> 28) << 12);
            const uint64_t has = __ballot(tmask != 0);
            if (has) {
                if (tmask) {
                    const uint32_t slot =
                        nrec + __builtin_amdgcn_mbcnt_hi((uint32_t)(has >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)has, 0u));
                    if (slot < qcap) {
                        uint32_t* rec = q_wave[slot];
                        rec[0] = lane | (k << 6) | (g << 9) | (min(gx, 31u) << 12);
                        rec[1] = tmask;
                        *reinterpret_cast<uint4*>(rec + 2) = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
                        // own bytes after the untied flips (tied ones are
                        // unchanged); packed, they are the row's stored bits
                        *reinterpret_cast<uint4*>(rec + 6) = make_uint4(n[0], n[1], n[2], n[3]);
                    } else {
                        // Queue full: resolve in this lane.  (A wave marching 8
                        // planes queues ~31 records on average against 64
                        // slots, so this happens a few times per 10^4
                        // sweeps of sc 256^3 x 3.)
                        uint32_t tm = tmask;
                        uint32_t flips = 0;
                        while (tm) {
                            const uint32_t j = __builtin_ctz(tm);
                            tm &= tm - 1u;
                            const uint32_t kk = j >> 2, sh = 8u * (j & 3u);
                            const uint32_t c = (pick4(kk, cnt[0], cnt[1], cnt[2], cnt[3]) >> sh) & 0xFFu;
                            const uint32_t bj = (pick4(kk, o[0], o[1], o[2], o[3]) >> sh) & 1u;
                            const uint32_t S = 2u * c + g + (j == gx ? 1u : 0u);
                            if (tie_accept(lut, bj ? always1 : always0, bj, S, i0 + j, row, t,
                                           colour, k0, k1)) {
                                if (!BITS) own_plane[off + j] = (uint8_t)(bj ^ 1u);
                                flips |= 1u << j;
                                const int32_t ds = 2 - 4 * (int32_t)bj;
                                m_part += ds;
                                if (colour == 1) b_part += ds * ((int32_t)S - 6);
                            }
                        }
                        if (BITS && flips)
                            __builtin_amdgcn_raw_buffer_store_b16(
                                (unsigned short)(pack16(n[0], n[1], n[2], n[3]) ^ flips), rw_own,
                                off, 0, 0);
                    }
                }
                nrec += __builtin_popcountll(has);
            }
            below = here;
            here = above;
            below_ok = true;
            off += zstep;
        }
    }

    // Resolve this wave's queued ties: one lane per record draws the low 24
    // bits of each tied site, writes the flipped byte and corrects the sums.
    // The wave's own row stores must land first.  nrec grows by ballot counts
    // among the lanes still marching, and a lane that stops (last z chunk)
    // or never started (past the chain's items, or the row y = ly of an odd
    // ly) stops counting: the lanes marching longest saw every record, so
    // the wave's maximum is the count, the same for every lane.
    const uint32_t nq = min(wave_max(nrec), qcap);
    if (nq) {
        __builtin_amdgcn_s_waitcnt(0);
        for (uint32_t e = lane; e < nq; e += 64) {
            const uint32_t* rec = q_wave[e];
            const uint32_t meta = rec[0];
            const uint32_t owner = (wave << 6) | (meta & 63u), k = (meta >> 6) & 7u;
            const uint32_t g = (meta >> 9) & 7u, gx = (meta >> 12) & 31u;
            const uint32_t l = (b - C.block0) * kBlock + owner;
            const uint32_t r = fast_div(l, C.segs_m, C.segs_s);
            const uint32_t sg = l - r * C.segs;
            const uint32_t r2 = fast_div(r, C.hy_m, C.hy_s);
            const uint32_t yy = 2u * (r - r2 * C.hy) + (r2 & 1u);
            const uint32_t zz = (r2 >> 1) * K + k;
            const uint32_t row = yy + ly * zz;
            uint32_t tm = rec[1];
            uint32_t flips = 0;
            while (tm) {
                const uint32_t j = __builtin_ctz(tm);
                tm &= tm - 1u;
                const uint32_t sh = 8u * (j & 3u);
                const uint32_t c = (rec[2 + (j >> 2)] >> sh) & 0xFFu;
                const uint32_t bj = (rec[6 + (j >> 2)] >> sh) & 1u;
                const uint32_t S = 2u * c + g + (j == gx ? 1u : 0u);
                const uint32_t pos = sg * kIsingSeg + j;
                if (tie_accept(lut, bj ? always1 : always0, bj, S, pos, row, t, colour, k0, k1)) {
                    if (!BITS) own_plane[row * pitch + pos] = (uint8_t)(bj ^ 1u);
                    flips |= 1u << j;
                    const int32_t ds = 2 - 4 * (int32_t)bj;
                    m_part += ds;
                    if (colour == 1) b_part += ds * ((int32_t)S - 6);
                }
            }
            // Bit-packed: the owner stored the row's 16 bits with the tied
            // sites unchanged; this lane rewrites them with the accepted ones.
            if (BITS && flips)
                __builtin_amdgcn_raw_buffer_store_b16(
                    (unsigned short)(pack16(rec[6], rec[7], rec[8], rec[9]) ^ flips), rw_own,
                    row * pitch + seg_off<BITS>(sg * kIsingSeg), 0, 0);
        }
    }

    m_part = wave_sum(m_part);
    b_part = wave_sum(b_part);
    if (lane == 0) {
        int32_t* p = partials + ((uint64_t)(s * 2u + colour) * total_blocks + b) * (2u * kWaves) +
                     2u * wave;
        p[0] = m_part;
        p[1] = b_part;
    }
}

template __global__ void ising_sc_pass<1, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<2, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<4, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<8, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<1, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<2, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<4, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);
template __global__ void ising_sc_pass<8, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t);

// Sums the pass partials of each (chain, sweep) over its ncol colour passes
// into its record: record = {bond, 0, 0, mag, 0} (int64).  A pass adds the
// bonds to lower colours only (none in colour 0), so each bond counts once.
// grid = (chains, sweeps).
__global__ __launch_bounds__(256) void ising_reduce(const ChainDev* __restrict__ chains,
                                                    const int32_t* __restrict__ partials,
                                                    uint32_t total_blocks, uint32_t ncol,
                                                    uint32_t nsw, int64_t* __restrict__ records) {
    __shared__ int64_t red[2][kBlock / 64];
    const ChainDev& C = chains[blockIdx.x];
    const uint32_t s = blockIdx.y;
    if (s >= C.limit) return;  // this chain ran fewer sweeps in the block
    int64_t mm = 0, bb = 0;
    for (uint32_t k = threadIdx.x; k < C.nblocks; k += kBlock) {
        const uint32_t blk = C.block0 + k;
        for (uint32_t c = 0; c < ncol; ++c) {
            const int32_t* p =
                partials + ((uint64_t)(s * ncol + c) * total_blocks + blk) * (2u * kWaves);
#pragma unroll
            for (uint32_t v = 0; v < kWaves; ++v) {
                mm += (int64_t)p[2 * v];
                bb += (int64_t)p[2 * v + 1];
            }
        }
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
