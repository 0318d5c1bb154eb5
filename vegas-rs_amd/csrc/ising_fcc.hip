// Int8 Ising colour pass on the face-centred cubic lattice, for any number of
// chains packed in one arena.
//
// The same update as the sc and bcc passes (MetropolisFlipIntegrator::step /
// MetropolisIntegrator::step, reference src/integrator.rs:66-138, with the
// Exchange + Zeeman local energy of src/energy.rs:147-151, 194-206) on the
// fcc topology of Lattice::fcc (src/input.rs:296-322), which is not
// bipartite: a sweep is four passes, one per sublattice (fcc_common.hpp), each
// reading the three other planes.
//
// Work item: one 16-position segment of a row, marched through K (the tile)
// consecutive z planes in the sc pass's work order.  For each other
// sublattice b the neighbour counts of the 16 sites are formed four sites
// per register; a and b differ along two axes:
//   y, z: the rows (y, z), (y', z), (y, z'), (y', z') at x, summed byte-wise;
//   x, z: the rows (y, z), (y, z'), summed, plus that sum shifted by a byte;
//   x, y: the rows (y, z), (y', z), summed, plus that sum shifted by a byte
// (the shift: x-1 for colours 0 and 1, x+1 for colours 2 and 3).  With z' =
// z - 1 or z + 1 fixed per pair, the rows of plane z' or z that the next
// plane needs again are kept in registers, so a plane loads its own row and
// five rows of the others instead of eight.  The decision is the bcc pass's,
// four sites per register (ising_swar.hpp): class (b, S) with S = 2 * (up
// neighbours) + (absent slots) of the 25-entry table, the 32-bit draw's high
// byte from one Philox call per 16 positions and its low 24 bits only on a
// high-byte tie; a row with ties leaves a record in the wave's LDS queue,
// settled once the wave has finished its planes.  The pass of colour a adds
// the bonds to the lower colours b < a, so over a sweep every bond counts
// once.  BITS: the planes hold one bit per site (ising_swar.hpp's storage
// forms); a segment is a 16-bit load expanded to the same 16 bytes.
#include "fcc_common.hpp"
#include "ising_swar.hpp"

namespace vg {

namespace {
constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kTieRows = 64;  // tie records per wave (one per row with ties)
constexpr uint32_t kTieWords = 14;
}  // namespace

template <uint32_t A, uint32_t K, bool BITS>
__global__ __launch_bounds__(256) void ising_fcc_pass(const ChainDev* __restrict__ chains,
                                                      const uint32_t* __restrict__ block_chain,
                                                      uint8_t* __restrict__ arena,
                                                      int32_t* __restrict__ partials, uint32_t s,
                                                      uint32_t total_blocks, uint32_t grid,
                                                      uint32_t opts) {
    __shared__ uint32_t lut[2 * kThrS];
    __shared__ uint32_t tab[13 * 8];  // class tables for 0..12 absent slots
    // per wave: kTieRows records of {meta, tie mask, cnt[4], own[4], low[4]}
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
    class_tables(C, 12u, tab);
    __syncthreads();

    const uint32_t local = (b - C.block0) * kBlock + tid;
    uint32_t (*const q_wave)[kTieWords] = tie_q[wave];
    uint32_t nrec = 0;  // records queued by this wave, as counted by this lane
    const uint64_t t = C.sweep0 + s;
    const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);
    const uint32_t always0 = C.always[0], always1 = C.always[1];
    uint8_t* const own_plane = arena + C.off[A];
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
        const uint32_t zstep = ly * pitch;
        const uint32_t i0 = seg * kIsingSeg;
        const uint32_t s0 = seg_off<BITS>(i0);  // the segment's byte offset in a row
        const uint32_t nvalid = min(kIsingSeg, w - i0);
        const uint32_t pw = w - i0;  // positions left in the row (1..16 in the last segment)
        const uint32_t vbits = nvalid >= kIsingSeg ? 0xFFFFu : (1u << nvalid) - 1u;
        uint32_t valid[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) valid[kk] = nibble_bytes(vbits >> (4 * kk)) << 7;

        // The x-offset neighbour outside the segment: position i0-1 (colours
        // 0, 1) or i0+16, or position 0 across the periodic end (colours 2, 3).
        constexpr bool kXneg = A < 2;
        const bool xr_in = i0 + kIsingSeg < w;
        const uint32_t xe = kXneg ? (i0 > 0 ? i0 - 1 : w - 1) : (xr_in ? i0 + kIsingSeg : 0);
        const bool xe_ok = kXneg ? (i0 > 0 || px) : (xr_in || px);
        // the one position whose x-offset neighbours are absent (open x)
        uint32_t gx = 0xFFFFFFFFu;
        if (!px) {
            if (kXneg && i0 == 0) gx = 0;
            if (!kXneg && !xr_in) gx = pw - 1;
        }

        // Byte sums of the 16 positions plus the x-offset shift of a row sum
        // P with e = the sum of the same rows at position xe.
        auto xshift = [&](const u32x4& P4, uint32_t e, uint32_t c4[4]) {
            const uint32_t P[4] = {P4.x, P4.y, P4.z, P4.w};
            uint32_t xs[4];
            if (kXneg) {
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
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) c4[kk] = P[kk] + xs[kk];
        };

        // Per other sublattice: the y' row, and the pair's z-plane state.
        // A pair that differs along z needs planes z and z' = z + sz; the
        // lower of the two is the upper of the plane before.
        const uint32_t z0 = zc * K;
        const uint32_t nk = min(K, lz - z0);
        uint32_t y2[4];
        bool yok[4];
        u32x4 lo_q[4];     // lower plane's row sum (y,z pairs: both rows; x,z pairs: one row)
        uint32_t lo_e[4];  // x,z pairs: the lower row's byte at xe
        bool lo_ok[4];
#pragma unroll
        for (uint32_t ob = 0; ob < 4; ++ob) {
            if (ob == A) continue;
            const __amdgpu_buffer_rsrc_t ro = plane_rsrc(arena + C.off[ob], (uint32_t)C.plane);
            y2[ob] = fcc_move(y, fcc_step(A, ob, 1), ly, py, yok[ob]);
            const uint32_t diff = fcc_diff(A, ob);
            if (!(diff & 4u)) continue;  // x,y pair: within plane z
            // The lower plane of the first plane z0: z0 - 1 when sz = -1, z0 when +1.
            const int sz = fcc_step(A, ob, 2);
            uint32_t zl;
            if (sz < 0) {
                lo_ok[ob] = z0 > 0 || pz;
                zl = z0 > 0 ? z0 - 1 : lz - 1;
            } else {
                lo_ok[ob] = true;
                zl = z0;
            }
            const uint32_t r0 = (y + ly * zl) * pitch;
            if (diff & 2u) {  // y, z
                const u32x4 a = load_seg<BITS>(ro, lo_ok[ob] ? r0 + s0 : kNoRow);
                const u32x4 c = load_seg<BITS>(
                    ro, lo_ok[ob] && yok[ob] ? (y2[ob] + ly * zl) * pitch + s0 : kNoRow);
                lo_q[ob] = a + c;
                lo_e[ob] = 0;
            } else {  // x, z
                lo_q[ob] = load_seg<BITS>(ro, lo_ok[ob] ? r0 + s0 : kNoRow);
                lo_e[ob] = load_pos<BITS>(ro, r0, xe, lo_ok[ob] && xe_ok);
            }
        }

        uint32_t off = (y + ly * z0) * pitch + s0;  // own row (y, z), this segment
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            if (k >= nk) break;
            // Falling priority while every wave is resident (as in the sc pass).
            if (prio) set_prio(((K - 1u - k) * 4u) / K);
            const uint32_t z = z0 + k;
            const uint32_t row = y + ly * z;
            const u32x4 o4 = load_seg<BITS>(rw_own, off);
            uint32_t cnt[4] = {0u, 0u, 0u, 0u}, low[4] = {0u, 0u, 0u, 0u};
            uint32_t pres = 0, pres_low = 0, miss = 0, miss_low = 0;
#pragma unroll
            for (uint32_t ob = 0; ob < 4; ++ob) {
                if (ob == A) continue;
                const uint32_t diff = fcc_diff(A, ob);
                const __amdgpu_buffer_rsrc_t ro = plane_rsrc(arena + C.off[ob], (uint32_t)C.plane);
                uint32_t c4[4], pr, ms = 0;
                if (diff & 4u) {
                    // the upper plane: z when sz = -1, z+1 when +1
                    const int sz = fcc_step(A, ob, 2);
                    bool hi_ok;
                    uint32_t zu;
                    if (sz < 0) {
                        hi_ok = true;
                        zu = z;
                    } else {
                        hi_ok = z + 1 < lz || pz;
                        zu = z + 1 < lz ? z + 1 : 0;
                    }
                    const uint32_t r0 = (y + ly * zu) * pitch;
                    if (diff & 2u) {  // y, z: four rows at x
                        const u32x4 a = load_seg<BITS>(ro, hi_ok ? r0 + s0 : kNoRow);
                        const u32x4 c = load_seg<BITS>(
                            ro, hi_ok && yok[ob] ? (y2[ob] + ly * zu) * pitch + s0 : kNoRow);
                        const u32x4 hq = a + c;
                        const u32x4 P = lo_q[ob] + hq;
                        c4[0] = P.x;
                        c4[1] = P.y;
                        c4[2] = P.z;
                        c4[3] = P.w;
                        pr = (lo_ok[ob] ? 1u : 0u) + (lo_ok[ob] && yok[ob] ? 1u : 0u) +
                             (hi_ok ? 1u : 0u) + (hi_ok && yok[ob] ? 1u : 0u);
                        lo_q[ob] = hq;
                    } else {  // x, z: two rows, each at x and at the x offset
                        const u32x4 hq = load_seg<BITS>(ro, hi_ok ? r0 + s0 : kNoRow);
                        const uint32_t he = load_pos<BITS>(ro, r0, xe, hi_ok && xe_ok);
                        xshift(lo_q[ob] + hq, lo_e[ob] + he, c4);
                        const uint32_t nr = (lo_ok[ob] ? 1u : 0u) + (hi_ok ? 1u : 0u);
                        pr = 2u * nr;
                        ms = nr;
                        lo_q[ob] = hq;
                        lo_e[ob] = he;
                    }
                    lo_ok[ob] = hi_ok;
                } else {
                    // x, y: rows (y, z) and (y', z), each at x and at the x offset
                    const uint32_t r1 = (y2[ob] + ly * z) * pitch;
                    const u32x4 v0 = load_seg<BITS>(ro, off);
                    const u32x4 v1 = load_seg<BITS>(ro, yok[ob] ? r1 + s0 : kNoRow);
                    const uint32_t e = load_pos<BITS>(ro, row * pitch, xe, xe_ok) +
                                       load_pos<BITS>(ro, r1, xe, xe_ok && yok[ob]);
                    xshift(v0 + v1, e, c4);
                    const uint32_t nr = yok[ob] ? 2u : 1u;
                    pr = 2u * nr;
                    ms = nr;
                }
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) cnt[kk] += c4[kk];  // 0..12 per byte
                pres += pr;
                miss += ms;
                if (ob < A) {
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) low[kk] += c4[kk];
                    pres_low += pr;
                    miss_low += ms;
                }
            }

            const uint32_t o[4] = {o4.x, o4.y, o4.z, o4.w};
            // The draws do not wait for the loads.  The compiler would sink
            // them past the branch below, behind the wait for the loads: an
            // empty asm statement taking them keeps them ahead of it.
            U4 ra = philox(U4{i0 >> 4, row, (uint32_t)t, ctr_w(t, kIsingA, A)}, k0, k1);
            asm volatile("" : "+v"(ra.x), "+v"(ra.y), "+v"(ra.z), "+v"(ra.w));
            const uint32_t rw[4] = {ra.x, ra.y, ra.z, ra.w};
            uint32_t live[4] = {valid[0], valid[1], valid[2], valid[3]};
            if (C.proposal) {
                const U4 rp =
                    philox(U4{i0 >> 7, row, (uint32_t)t, ctr_w(t, kIsingProposal, A)}, k0, k1);
                const uint32_t bits = (u4_word(rp, (i0 >> 5) & 3u) >> (i0 & 31u)) & 0xFFFFu;
                // Spin::rand proposes Up on bit 0 (state.rs:76-84); a proposal
                // equal to the current spin changes nothing.
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    live[kk] &= (~(nibble_bytes(bits >> (4 * kk)) ^ o[kk]) & kOne) << 7;
            }
            // absent slots of the segment's sites, and of the position gx
            const uint32_t a_in = 12u - pres, a_edge = a_in + miss;
            uint32_t T0[4], T1[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                T0[kk] = tab[a_in * 8u + kk];
                T1[kk] = tab[a_in * 8u + 4u + kk];
            }
            uint32_t n[4], accH[4], tieH[4];
            decide16(cnt, o, rw, live, T0, T1, n, accH, tieH);
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
            uint32_t nl = 0, ls = 0;
            int32_t upc = 0;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint32_t lv = low[kk] & ((valid[kk] >> 7) * 0xFFu);
                nl = __builtin_amdgcn_udot4(n[kk], lv, nl, false);
                ls = __builtin_amdgcn_udot4(lv, kOne, ls, false);
                upc += __popc(n[kk]);
            }
            m_part += 2 * upc - (int32_t)nvalid;
            if (A > 0) {
                // bonds to the lower colours: sum_j s_j (2 low_j - pl_j) with
                // s = 2n - 1 and pl the present lower slots (fewer by
                // miss_low at gx)
                b_part += 4 * (int32_t)nl - 2 * (int32_t)ls -
                          (int32_t)pres_low * (2 * upc - (int32_t)nvalid);
                if (gx < nvalid) {
                    const int32_t nj = (int32_t)((n[gx >> 2] >> (8u * (gx & 3u))) & 1u);
                    b_part += (2 * nj - 1) * (int32_t)miss_low;
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
                                 (min(gx, 31u) << 17) | (pres_low << 22) | (miss_low << 26);
                        rec[1] = tmask;
                        *reinterpret_cast<uint4*>(rec + 2) = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
                        *reinterpret_cast<uint4*>(rec + 6) = make_uint4(n[0], n[1], n[2], n[3]);
                        *reinterpret_cast<uint4*>(rec + 10) = make_uint4(low[0], low[1], low[2], low[3]);
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
                            if (tie_accept(lut, bj ? always1 : always0, bj, S, i0 + j, row, t, A,
                                           k0, k1)) {
                                if (!BITS) own_plane[off + j] = (uint8_t)(bj ^ 1u);
                                flips |= 1u << j;
                                const int32_t ds = 2 - 4 * (int32_t)bj;
                                m_part += ds;
                                if (A > 0) {
                                    const uint32_t pl = pres_low - (j == gx ? miss_low : 0u);
                                    b_part += ds * (2 * (int32_t)((low[kk] >> sh) & 0xFFu) -
                                                    (int32_t)pl);
                                }
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
            const uint32_t pres_low = (meta >> 22) & 15u, miss_low = (meta >> 26) & 7u;
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
                if (tie_accept(lut, bj ? always1 : always0, bj, S, pos, row, t, A, k0, k1)) {
                    if (!BITS) own_plane[row * pitch + pos] = (uint8_t)(bj ^ 1u);
                    flips |= 1u << j;
                    const int32_t ds = 2 - 4 * (int32_t)bj;
                    m_part += ds;
                    if (A > 0) {
                        const uint32_t lw = (rec[10 + (j >> 2)] >> sh) & 0xFFu;
                        const uint32_t pl = pres_low - (j == gx ? miss_low : 0u);
                        b_part += ds * (2 * (int32_t)lw - (int32_t)pl);
                    }
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
        int32_t* p = partials + ((uint64_t)(s * 4u + A) * total_blocks + b) * (2u * kWaves) +
                     2u * wave;
        p[0] = m_part;
        p[1] = b_part;
    }
}

#define VG_FCC_PASS(A, K)                                                                    \
    template __global__ void ising_fcc_pass<A, K, false>(const ChainDev*, const uint32_t*,     \
                                                         uint8_t*, int32_t*, uint32_t, uint32_t, \
                                                         uint32_t, uint32_t);                    \
    template __global__ void ising_fcc_pass<A, K, true>(const ChainDev*, const uint32_t*,      \
                                                        uint8_t*, int32_t*, uint32_t, uint32_t,  \
                                                        uint32_t, uint32_t);
VG_FCC_PASS(0, 1)
VG_FCC_PASS(0, 2)
VG_FCC_PASS(0, 4)
VG_FCC_PASS(0, 8)
VG_FCC_PASS(1, 1)
VG_FCC_PASS(1, 2)
VG_FCC_PASS(1, 4)
VG_FCC_PASS(1, 8)
VG_FCC_PASS(2, 1)
VG_FCC_PASS(2, 2)
VG_FCC_PASS(2, 4)
VG_FCC_PASS(2, 8)
VG_FCC_PASS(3, 1)
VG_FCC_PASS(3, 2)
VG_FCC_PASS(3, 4)
VG_FCC_PASS(3, 8)
#undef VG_FCC_PASS

}  // namespace vg
