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
// Work item: one 16-position segment of a row.  For each other sublattice b
// the neighbour counts of the 16 sites are formed four sites per register:
// when a and b differ along y and z, the four neighbour rows at the same x
// summed byte-wise; when they differ along x and one of y, z, the two rows
// summed and that sum plus itself shifted by one byte (x-1 for colours 0
// and 1, x+1 for colours 2 and 3).  The decision is the bcc pass's, four
// sites per register (ising_swar.hpp): class (b, S) with S = 2 * (up
// neighbours) + (absent slots) of the 25-entry table, the 32-bit draw's high
// byte from one Philox call per 16 positions and its low 24 bits only on a
// high-byte tie.  The pass of colour a adds the bonds to the lower colours
// b < a, so over a sweep every bond counts once.
#include "fcc_common.hpp"
#include "ising_swar.hpp"

namespace vg {

namespace {
constexpr uint32_t kWaves = kBlock / 64;
}

template <uint32_t A>
__global__ __launch_bounds__(256) void ising_fcc_pass(const ChainDev* __restrict__ chains,
                                                      const uint32_t* __restrict__ block_chain,
                                                      uint8_t* __restrict__ arena,
                                                      int32_t* __restrict__ partials, uint32_t s,
                                                      uint32_t total_blocks, uint32_t grid) {
    __shared__ uint32_t lut[2 * kThrS];
    __shared__ uint32_t tab[13 * 8];  // class tables for 0..12 absent slots
    const uint32_t b = xcd_block(blockIdx.x, grid);
    if (b >= total_blocks) return;  // uniform over the workgroup
    const ChainDev& C = chains[block_chain[b]];
    if (s >= C.limit) return;  // uniform: the workgroup belongs to one chain
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid < 2 * kThrS) lut[tid] = (&C.thr[0][0])[tid];
    class_tables(C, 12u, tab);
    __syncthreads();

    const uint32_t local = (b - C.block0) * kBlock + tid;
    int32_t m_part = 0, b_part = 0;
    if (local < C.items) {
        const uint64_t t = C.sweep0 + s;
        const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);
        const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
        const uint32_t row = fast_div(local, C.segs_m, C.segs_s);
        const uint32_t seg = local - row * C.segs;
        const uint32_t z = fast_div(row, C.ly_m, C.ly_s);
        const uint32_t y = row - z * ly;
        const bool px = pbc & 1u, py = pbc & 2u, pz = pbc & 4u;
        uint8_t* const own_plane = arena + C.off[A];
        const __amdgpu_buffer_rsrc_t rw_own = plane_rsrc(own_plane, (uint32_t)C.plane);

        const uint32_t i0 = seg * kIsingSeg;
        const uint32_t nvalid = min(kIsingSeg, w - i0);
        const uint32_t pw = w - i0;  // positions left in the row (1..16 in the last segment)
        const uint32_t off = row * pitch + i0;
        const u32x4 o4 = bload16(rw_own, off);

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

        uint32_t cnt[4] = {0u, 0u, 0u, 0u}, low[4] = {0u, 0u, 0u, 0u};
        uint32_t pres = 0, pres_low = 0, miss = 0, miss_low = 0;
#pragma unroll
        for (uint32_t ob = 0; ob < 4; ++ob) {
            if (ob == A) continue;
            const uint32_t diff = fcc_diff(A, ob);
            const __amdgpu_buffer_rsrc_t ro = plane_rsrc(arena + C.off[ob], (uint32_t)C.plane);
            bool yok, zok;
            const uint32_t y2 = fcc_move(y, fcc_step(A, ob, 1), ly, py, yok);
            const uint32_t z2 = fcc_move(z, fcc_step(A, ob, 2), lz, pz, zok);
            uint32_t c4[4], pr, ms = 0;
            if (!(diff & 1u)) {
                // y and z differ: rows (y,z), (y',z), (y,z'), (y',z') at x
                const uint32_t rows[4] = {row, y2 + ly * z, y + ly * z2, y2 + ly * z2};
                const bool rok[4] = {true, yok, zok, yok && zok};
                c4[0] = c4[1] = c4[2] = c4[3] = 0u;
                pr = 0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const u32x4 v = bload16(ro, rok[r] ? rows[r] * pitch + i0 : kNoRow);
                    c4[0] += v.x;
                    c4[1] += v.y;
                    c4[2] += v.z;
                    c4[3] += v.w;
                    pr += rok[r] ? 1u : 0u;
                }
            } else {
                // x and one of y, z differ: rows (y,z) and the shifted one,
                // each at x and at the x-offset position
                const bool ok2 = (diff & 2u) ? yok : zok;
                const uint32_t row2 = (diff & 2u) ? y2 + ly * z : y + ly * z2;
                const u32x4 v0 = bload16(ro, off);
                const u32x4 v1 = bload16(ro, ok2 ? row2 * pitch + i0 : kNoRow);
                uint32_t P[4] = {v0.x + v1.x, v0.y + v1.y, v0.z + v1.z, v0.w + v1.w};
                uint32_t e = bload8(ro, xe_ok ? row * pitch + xe : kNoRow) +
                             bload8(ro, xe_ok && ok2 ? row2 * pitch + xe : kNoRow);
                uint32_t xs[4];
                if (kXneg) {
                    xs[0] = __builtin_amdgcn_alignbyte(P[0], e << 24, 3);
                    xs[1] = __builtin_amdgcn_alignbyte(P[1], P[0], 3);
                    xs[2] = __builtin_amdgcn_alignbyte(P[2], P[1], 3);
                    xs[3] = __builtin_amdgcn_alignbyte(P[3], P[2], 3);
                } else {
                    uint32_t m[4] = {P[0], P[1], P[2], P[3]};
                    if (!xr_in && pw != kIsingSeg) {
                        // byte pw (padding, zero) takes the wrapped neighbour
                        // of position w-1
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
                const uint32_t nr = ok2 ? 2u : 1u;
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
        const U4 ra = philox(U4{i0 >> 4, row, (uint32_t)t, ctr_w(t, kIsingA, A)}, k0, k1);
        const uint32_t rw[4] = {ra.x, ra.y, ra.z, ra.w};
        const uint32_t vbits = nvalid >= kIsingSeg ? 0xFFFFu : (1u << nvalid) - 1u;
        uint32_t valid[4], live[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) live[kk] = valid[kk] = nibble_bytes(vbits >> (4 * kk)) << 7;
        if (C.proposal) {
            const U4 rp = philox(U4{i0 >> 7, row, (uint32_t)t, ctr_w(t, kIsingProposal, A)}, k0, k1);
            const uint32_t bits = (u4_word(rp, (i0 >> 5) & 3u) >> (i0 & 31u)) & 0xFFFFu;
            // Spin::rand proposes Up on bit 0 (state.rs:76-84); a proposal
            // equal to the current spin changes nothing.
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                live[kk] &= (~(nibble_bytes(bits >> (4 * kk)) ^ o[kk]) & kOne) << 7;
        }
        // absent slots of the segment's sites, and of the position gx
        const uint32_t a_in = 12u - pres, a_edge = a_in + miss;
        const uint32_t always0 = C.always[0], always1 = C.always[1];
        uint32_t T0[4], T1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            T0[k] = tab[a_in * 8u + k];
            T1[k] = tab[a_in * 8u + 4u + k];
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
        // Tied sites: the low 24 bits of the draw decide.
        uint32_t tm = msb_mask16(tieH);
        while (tm) {
            const uint32_t j = __builtin_ctz(tm);
            tm &= tm - 1u;
            const uint32_t kk = j >> 2, sh = 8u * (j & 3u);
            const uint32_t bj = (o[kk] >> sh) & 1u;
            const uint32_t S = 2u * ((cnt[kk] >> sh) & 0xFFu) + (j == gx ? a_edge : a_in);
            if (tie_accept(lut, bj ? always1 : always0, bj, S, i0 + j, row, t, A, k0, k1))
                n[kk] ^= 1u << sh;
        }
        if ((n[0] ^ o[0]) | (n[1] ^ o[1]) | (n[2] ^ o[2]) | (n[3] ^ o[3])) {
            u32x4 nv;
            nv.x = n[0];
            nv.y = n[1];
            nv.z = n[2];
            nv.w = n[3];
            __builtin_amdgcn_raw_buffer_store_b128(nv, rw_own, off, 0, 0);
        }
        if (A > 0) {
            // bonds to the lower colours: sum_j s_j (2 low_j - pl_j) with
            // s = 2n - 1 and pl the present lower slots (fewer by miss_low at gx)
            uint32_t nl = 0, ls = 0;
            int32_t upc = 0;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint32_t lv = low[kk] & ((valid[kk] >> 7) * 0xFFu);
                nl = __builtin_amdgcn_udot4(n[kk], lv, nl, false);
                ls = __builtin_amdgcn_udot4(lv, kOne, ls, false);
                upc += __popc(n[kk]);
            }
            b_part = 4 * (int32_t)nl - 2 * (int32_t)ls -
                     (int32_t)pres_low * (2 * upc - (int32_t)nvalid);
            if (gx < nvalid) {
                const int32_t nj = (int32_t)((n[gx >> 2] >> (8u * (gx & 3u))) & 1u);
                b_part += (2 * nj - 1) * (int32_t)miss_low;
            }
        }
        const int32_t up = __popc(n[0]) + __popc(n[1]) + __popc(n[2]) + __popc(n[3]);
        m_part = 2 * up - (int32_t)nvalid;
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

template __global__ void ising_fcc_pass<0>(const ChainDev*, const uint32_t*, uint8_t*, int32_t*,
                                           uint32_t, uint32_t, uint32_t);
template __global__ void ising_fcc_pass<1>(const ChainDev*, const uint32_t*, uint8_t*, int32_t*,
                                           uint32_t, uint32_t, uint32_t);
template __global__ void ising_fcc_pass<2>(const ChainDev*, const uint32_t*, uint8_t*, int32_t*,
                                           uint32_t, uint32_t, uint32_t);
template __global__ void ising_fcc_pass<3>(const ChainDev*, const uint32_t*, uint8_t*, int32_t*,
                                           uint32_t, uint32_t, uint32_t);

}  // namespace vg
