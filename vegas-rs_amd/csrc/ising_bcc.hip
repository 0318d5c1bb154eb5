// Int8 Ising checkerboard colour pass on the body-centred cubic lattice, for
// any number of chains packed in one arena.
//
// The same update as the simple-cubic pass (MetropolisFlipIntegrator::step /
// MetropolisIntegrator::step, reference src/integrator.rs:66-138, with the
// Exchange + Zeeman local energy of src/energy.rs:147-151, 194-206) on the
// bcc topology: the two colours are the two sublattices.  A corner site of
// cell (x,y,z) (colour 0) bonds to the body centres of cells (x-dx, y-dy,
// z-dz), a body centre (colour 1) to the corners of cells (x+dx, y+dy, z+dz),
// dx,dy,dz in {0,1}: four rows of the other plane, each at positions x and
// x-1 (colour 0) or x and x+1 (colour 1).
//
// Work item: one 16-position segment of a row, marched through K (the tile)
// consecutive z planes, in the sc pass's work order.  The four neighbour rows
// of plane z are two pairs, (y, zl), (y', zl) and (y, zu), (y', zu), with
// (zl, zu) = (z-1, z) for colour 0 and (z, z+1) for colour 1; marching up,
// the upper pair of plane z is the lower pair of plane z+1, so a plane loads
// its own row and one pair (three 16-byte loads and two bytes) instead of
// five rows and four bytes.  The neighbour counts of the 16 sites are formed
// four sites per register (the pair sums added byte-wise, then that sum plus
// itself shifted by one byte), and so is the decision (ising_swar.hpp): class
// (b, S) with S = 2 * (up neighbours) + (absent slots), accepted when always,
// or when the 32-bit draw (high byte from one Philox call per 16 positions,
// low 24 bits drawn only on a high-byte tie) is below the class threshold.
// As in the sc pass, a row with ties leaves a record in the wave's LDS queue
// and one lane per record settles them once the wave has finished its planes.
// BITS: the planes hold one bit per site (ising_swar.hpp's storage forms);
// a segment is a 16-bit load expanded to the same 16 bytes.
#include "ising_swar.hpp"

namespace vg {

namespace {
constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kTieRows = 64;  // tie records per wave (one per row with ties)
constexpr uint32_t kTieWords = 10;
}

template <uint32_t K, bool BITS>
__global__ __launch_bounds__(256) void ising_bcc_pass(const ChainDev* __restrict__ chains,
                                                      const uint32_t* __restrict__ block_chain,
                                                      uint8_t* __restrict__ arena,
                                                      int32_t* __restrict__ partials,
                                                      uint32_t colour, uint32_t s,
                                                      uint32_t total_blocks, uint32_t grid,
                                                      uint32_t opts) {
    __shared__ uint32_t lut[2 * kThrS];
    __shared__ uint32_t tab[9 * 8];  // class tables for 0..8 absent slots
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
    class_tables(C, 8u, tab);
    __syncthreads();

    const uint32_t local = (b - C.block0) * kBlock + tid;
    uint32_t (*const q_wave)[kTieWords] = tie_q[wave];
    uint32_t nrec = 0;  // records queued by this wave, as counted by this lane
    const uint64_t t = C.sweep0 + s;
    const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);
    const uint32_t always0 = C.always[0], always1 = C.always[1];
    uint8_t* const own_plane = arena + C.off[colour];
    const __amdgpu_buffer_rsrc_t rw_own = plane_rsrc(own_plane, (uint32_t)C.plane);
    int32_t m_part = 0, b_part = 0;
    uint32_t seg = 0, y = 0, zc = 0;
    if (local < C.items) {
        const uint32_t r = fast_div(local, C.segs_m, C.segs_s);
        seg = local - r * C.segs;
        const uint32_t r2 = fast_div(r, C.hy_m, C.hy_s);
        y = 2u * (r - r2 * C.hy) + (r2 & 1u);
        zc = r2 >> 1;
    }
    if (local < C.items && y < C.ly) {
        const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
        const bool px = pbc & 1u, py = pbc & 2u, pz = pbc & 4u;
        const __amdgpu_buffer_rsrc_t ro = plane_rsrc(arena + C.off[colour ^ 1u], (uint32_t)C.plane);
        const uint32_t zstep = ly * pitch;

        // y' = y - 1 (colour 0) or y + 1 (colour 1), wrapping when periodic
        bool yok;
        uint32_t y2;
        if (colour == 0) {
            yok = y > 0 || py;
            y2 = y > 0 ? y - 1 : ly - 1;
        } else {
            yok = y + 1 < ly || py;
            y2 = y + 1 < ly ? y + 1 : 0;
        }
        const uint32_t i0 = seg * kIsingSeg;
        const uint32_t nvalid = min(kIsingSeg, w - i0);
        const uint32_t vbits = nvalid >= kIsingSeg ? 0xFFFFu : (1u << nvalid) - 1u;
        uint32_t valid[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) valid[kk] = nibble_bytes(vbits >> (4 * kk)) << 7;
        // The x-offset position outside the segment: x-1 (colour 0) or x+1
        // (colour 1; position 0 across the periodic end).  gx: the one
        // position whose x-offset neighbours are absent (open x).
        bool x_ok;
        uint32_t xpos, gx = 0xFFFFFFFFu;
        const uint32_t pw = w - i0;  // positions left in the row, 1..16 in the last segment
        const bool xr_in = i0 + kIsingSeg < w;
        if (colour == 0) {
            x_ok = i0 > 0 || px;
            xpos = i0 > 0 ? i0 - 1 : w - 1;
            if (!x_ok) gx = 0;
        } else {
            x_ok = xr_in || px;
            xpos = xr_in ? i0 + kIsingSeg : 0;
            if (!xr_in && !px) gx = pw - 1;
        }
        // The pair of rows (y, zz), (y', zz): byte sums at the segment (q) and
        // at xpos (e); absent rows read zeros.
        const uint32_t s0 = seg_off<BITS>(i0);
        auto load_pair = [&](uint32_t zz, bool zok, u32x4& q, uint32_t& e) {
            const uint32_t r0 = (y + ly * zz) * pitch, r1 = (y2 + ly * zz) * pitch;
            const u32x4 a = load_seg<BITS>(ro, zok ? r0 + s0 : kNoRow);
            const u32x4 c = load_seg<BITS>(ro, zok && yok ? r1 + s0 : kNoRow);
            q = a + c;
            e = load_pos<BITS>(ro, r0, xpos, zok && x_ok) +
                load_pos<BITS>(ro, r1, xpos, zok && x_ok && yok);
        };

        const uint32_t z0 = zc * K;
        const uint32_t nk = min(K, lz - z0);
        // The lower pair of the first plane: zl = z0-1 (colour 0) or z0.
        u32x4 q_lo;
        uint32_t e_lo;
        bool lo_ok;
        if (colour == 0) {
            lo_ok = z0 > 0 || pz;
            load_pair(z0 > 0 ? z0 - 1 : lz - 1, lo_ok, q_lo, e_lo);
        } else {
            lo_ok = true;
            load_pair(z0, true, q_lo, e_lo);
        }
        uint32_t off = (y + ly * z0) * pitch + s0;  // own row (y, z), this segment
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            if (k >= nk) break;
            // Falling priority while every wave is resident (as in the sc pass).
            if (prio) set_prio(((K - 1u - k) * 4u) / K);
            const uint32_t z = z0 + k;
            const uint32_t row = y + ly * z;
            // the upper pair: zu = z (colour 0) or z+1 (colour 1)
            bool hi_ok;
            uint32_t zu;
            if (colour == 0) {
                hi_ok = true;
                zu = z;
            } else {
                hi_ok = z + 1 < lz || pz;
                zu = z + 1 < lz ? z + 1 : 0;
            }
            u32x4 q_hi;
            uint32_t e_hi;
            load_pair(zu, hi_ok, q_hi, e_hi);
            const u32x4 o4 = load_seg<BITS>(rw_own, off);
            // The draws do not wait for the loads.  The compiler would sink
            // them past the branches below, behind the wait for the loads:
            // an empty asm statement taking them keeps them ahead.
            U4 ra = philox(U4{i0 >> 4, row, (uint32_t)t, ctr_w(t, kIsingA, colour)}, k0, k1);
            asm volatile("" : "+v"(ra.x), "+v"(ra.y), "+v"(ra.z), "+v"(ra.w));
            // present rows of the four; absent slots 8 - 2 nr (8 - nr at gx)
            const uint32_t nr = (lo_ok ? 1u : 0u) + (lo_ok && yok ? 1u : 0u) + (hi_ok ? 1u : 0u) +
                                (hi_ok && yok ? 1u : 0u);
            const u32x4 P4 = q_lo + q_hi;
            uint32_t e = e_lo + e_hi;
            const uint32_t P[4] = {P4.x, P4.y, P4.z, P4.w};
            uint32_t xs[4];
            if (colour == 0) {
                xs[0] = __builtin_amdgcn_alignbyte(P[0], e << 24, 3);
                xs[1] = __builtin_amdgcn_alignbyte(P[1], P[0], 3);
                xs[2] = __builtin_amdgcn_alignbyte(P[2], P[1], 3);
                xs[3] = __builtin_amdgcn_alignbyte(P[3], P[2], 3);
            } else {
                uint32_t m[4] = {P[0], P[1], P[2], P[3]};
                if (!xr_in && pw != kIsingSeg) {
                    // byte pw (padding, zero) takes the wrapped neighbour of
                    // position w-1
#pragma unroll
                    for (uint32_t kk = 0; kk < 4; ++kk)
                        if ((pw >> 2) == kk) m[kk] |= e << (8u * (pw & 3u));
                    e = 0;
                }
                xs[0] = __builtin_amdgcn_alignbyte(m[1], m[0], 1);
                xs[1] = __builtin_amdgcn_alignbyte(m[2], m[1], 1);
                xs[2] = __builtin_amdgcn_alignbyte(m[3], m[2], 1);
                xs[3] = __builtin_amdgcn_alignbyte(e, m[3], 1);
            }
            uint32_t cnt[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) cnt[kk] = P[kk] + xs[kk];  // 0..8 per byte

            const uint32_t o[4] = {o4.x, o4.y, o4.z, o4.w};
            const uint32_t rw[4] = {ra.x, ra.y, ra.z, ra.w};
            uint32_t live[4] = {valid[0], valid[1], valid[2], valid[3]};
            if (C.proposal) {
                const U4 rp = philox(
                    U4{i0 >> 7, row, (uint32_t)t, ctr_w(t, kIsingProposal, colour)}, k0, k1);
                const uint32_t bits = (u4_word(rp, (i0 >> 5) & 3u) >> (i0 & 31u)) & 0xFFFFu;
                // Spin::rand proposes Up on bit 0 (state.rs:76-84); a proposal
                // equal to the current spin changes nothing.
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    live[kk] &= (~(nibble_bytes(bits >> (4 * kk)) ^ o[kk]) & kOne) << 7;
            }
            const uint32_t a_in = 8u - 2u * nr, a_edge = 8u - nr;
            uint32_t T0[4], T1[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                T0[kk] = tab[a_in * 8u + kk];
                T1[kk] = tab[a_in * 8u + 4u + kk];
            }
            uint32_t n[4], accH[4], tieH[4];
            decide16(cnt, o, rw, live, T0, T1, n, accH, tieH);
            // The position with absent x-offset neighbours has S = 2c + a_edge.
            if (gx < nvalid) {
                const uint32_t kk = gx >> 2, sh = 8u * (gx & 3u);
                const uint32_t c = (cnt[kk] >> sh) & 0xFFu, bj = (o[kk] >> sh) & 1u;
                const uint32_t S = 2u * c + a_edge;
                const bool alw = ((bj ? always1 : always0) >> S) & 1u;
                const uint32_t th = alw ? 0xFFu : lut[bj * kThrS + S] >> 24;
                const uint32_t r8 = (rw[kk] >> sh) & 0xFFu;
                const bool lv = (live[kk] >> sh) & 0x80u;
                const uint32_t f = lv && r8 < th ? 1u : 0u;
                tieH[kk] = (tieH[kk] & ~(0x80u << sh)) | ((lv && r8 == th ? 0x80u : 0u) << sh);
                n[kk] = (n[kk] & ~(1u << sh)) | ((bj ^ f) << sh);
            }
            if ((n[0] ^ o[0]) | (n[1] ^ o[1]) | (n[2] ^ o[2]) | (n[3] ^ o[3])) {
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
            uint32_t nc = 0, cs = 0;
            int32_t upc = 0;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint32_t cv = cnt[kk] & ((valid[kk] >> 7) * 0xFFu);
                nc = __builtin_amdgcn_udot4(n[kk], cv, nc, false);
                cs = __builtin_amdgcn_udot4(cv, kOne, cs, false);
                upc += __popc(n[kk]);
            }
            m_part += 2 * upc - (int32_t)nvalid;
            if (colour == 1) {
                // bond sum over colour-1 sites: sum_j s_j (S_j - 8) with
                // s = 2n - 1 and S = 2c + a (a_edge at gx)
                b_part += 4 * (int32_t)nc - 2 * (int32_t)cs +
                          ((int32_t)a_in - 8) * (2 * upc - (int32_t)nvalid);
                if (gx < nvalid) {
                    const int32_t nj = (int32_t)((n[gx >> 2] >> (8u * (gx & 3u))) & 1u);
                    b_part += (2 * nj - 1) * ((int32_t)a_edge - (int32_t)a_in);
                }
            }
            // Rows with ties leave one record each in the wave's queue (their
            // tied bytes were written unchanged above); slots by ballot.
            const uint32_t tmask = msb_mask16(tieH);
            const uint64_t has = __ballot(tmask != 0);
            if (has) {
                if (tmask) {
                    const uint32_t slot =
                        nrec + __builtin_amdgcn_mbcnt_hi((uint32_t)(has >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)has, 0u));
                    if (slot < qcap) {
                        uint32_t* rec = q_wave[slot];
                        rec[0] = lane | (k << 6) | (a_in << 9) | (a_edge << 13) |
                                 (min(gx, 31u) << 17);
                        rec[1] = tmask;
                        *reinterpret_cast<uint4*>(rec + 2) = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
                        *reinterpret_cast<uint4*>(rec + 6) = make_uint4(n[0], n[1], n[2], n[3]);
                    } else {
                        // Queue full: resolve in this lane.  (A wave marching 8
                        // planes queues ~31 records on average against 64
                        // slots, so this happens a few times per 10^4
                        // sweeps of sc 256^3 x 3.)
                        uint32_t tm = tmask, flips = 0;
                        while (tm) {
                            const uint32_t j = __builtin_ctz(tm);
                            tm &= tm - 1u;
                            const uint32_t kk = j >> 2, sh = 8u * (j & 3u);
                            const uint32_t bj = (o[kk] >> sh) & 1u;
                            const uint32_t S =
                                2u * ((cnt[kk] >> sh) & 0xFFu) + (j == gx ? a_edge : a_in);
                            if (tie_accept(lut, bj ? always1 : always0, bj, S, i0 + j, row, t,
                                           colour, k0, k1)) {
                                if (!BITS) own_plane[off + j] = (uint8_t)(bj ^ 1u);
                                flips |= 1u << j;
                                const int32_t ds = 2 - 4 * (int32_t)bj;
                                m_part += ds;
                                if (colour == 1) b_part += ds * ((int32_t)S - 8);
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
            q_lo = q_hi;
            e_lo = e_hi;
            lo_ok = hi_ok;
            off += zstep;
        }
    }

    // Resolve this wave's queued ties: one lane per record draws the low 24
    // bits of each tied site, writes the flipped byte and corrects the sums.
    // The wave's own row stores must land first.  A lane stops counting
    // records when it stops marching (last z chunk) or never starts, so the
    // wave's maximum is the count.
    const uint32_t nq = min(wave_max(nrec), qcap);
    if (nq) {
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t ly = C.ly, pitch = C.pitch;
        for (uint32_t e = lane; e < nq; e += 64) {
            const uint32_t* rec = q_wave[e];
            const uint32_t meta = rec[0];
            const uint32_t owner = (wave << 6) | (meta & 63u), k = (meta >> 6) & 7u;
            const uint32_t a_in = (meta >> 9) & 15u, a_edge = (meta >> 13) & 15u;
            const uint32_t gx = (meta >> 17) & 31u;
            const uint32_t l = (b - C.block0) * kBlock + owner;
            const uint32_t r = fast_div(l, C.segs_m, C.segs_s);
            const uint32_t sg = l - r * C.segs;
            const uint32_t r2 = fast_div(r, C.hy_m, C.hy_s);
            const uint32_t yy = 2u * (r - r2 * C.hy) + (r2 & 1u);
            const uint32_t zz = (r2 >> 1) * K + k;
            const uint32_t row = yy + ly * zz;
            uint32_t tm = rec[1], flips = 0;
            while (tm) {
                const uint32_t j = __builtin_ctz(tm);
                tm &= tm - 1u;
                const uint32_t sh = 8u * (j & 3u);
                const uint32_t c = (rec[2 + (j >> 2)] >> sh) & 0xFFu;
                const uint32_t bj = (rec[6 + (j >> 2)] >> sh) & 1u;
                const uint32_t S = 2u * c + (j == gx ? a_edge : a_in);
                const uint32_t pos = sg * kIsingSeg + j;
                if (tie_accept(lut, bj ? always1 : always0, bj, S, pos, row, t, colour, k0, k1)) {
                    if (!BITS) own_plane[row * pitch + pos] = (uint8_t)(bj ^ 1u);
                    flips |= 1u << j;
                    const int32_t ds = 2 - 4 * (int32_t)bj;
                    m_part += ds;
                    if (colour == 1) b_part += ds * ((int32_t)S - 8);
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

template __global__ void ising_bcc_pass<1, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<2, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<4, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<8, false>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<1, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<2, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<4, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);
template __global__ void ising_bcc_pass<8, true>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t);

}  // namespace vg
