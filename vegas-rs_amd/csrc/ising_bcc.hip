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
// Work item: one 16-position segment of a row.  The neighbour counts of the
// 16 sites are formed four sites per register (the four rows summed
// byte-wise, then that sum plus itself shifted by one byte), and so is the
// decision (ising_swar.hpp): class (b, S) with S = 2 * (up neighbours) +
// (absent slots), accepted when always, or when the 32-bit draw (high byte
// from one Philox call per 16 positions, low 24 bits drawn only on a
// high-byte tie) is below the class threshold.
#include "ising_swar.hpp"

namespace vg {

namespace {
constexpr uint32_t kWaves = kBlock / 64;
}

__global__ __launch_bounds__(256) void ising_bcc_pass(const ChainDev* __restrict__ chains,
                                                      const uint32_t* __restrict__ block_chain,
                                                      uint8_t* __restrict__ arena,
                                                      int32_t* __restrict__ partials,
                                                      uint32_t colour, uint32_t s,
                                                      uint32_t total_blocks, uint32_t grid) {
    __shared__ uint32_t lut[2 * kThrS];
    __shared__ uint32_t tab[9 * 8];  // class tables for 0..8 absent slots
    const uint32_t b = xcd_block(blockIdx.x, grid);
    if (b >= total_blocks) return;  // uniform over the workgroup
    const ChainDev& C = chains[block_chain[b]];
    if (s >= C.limit) return;  // uniform: the workgroup belongs to one chain
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid < 2 * kThrS) lut[tid] = (&C.thr[0][0])[tid];
    class_tables(C, 8u, tab);
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
        uint8_t* const own_plane = arena + C.off[colour];
        const __amdgpu_buffer_rsrc_t rw_own = plane_rsrc(own_plane, (uint32_t)C.plane);
        const __amdgpu_buffer_rsrc_t ro = plane_rsrc(arena + C.off[colour ^ 1u], (uint32_t)C.plane);

        // The neighbour rows: (y, z), (y', z), (y, z'), (y', z') with
        // y' = y - 1, z' = z - 1 (colour 0) or y + 1, z + 1 (colour 1).
        bool yok, zok;
        uint32_t y2, z2;
        if (colour == 0) {
            yok = y > 0 || py;
            y2 = y > 0 ? y - 1 : ly - 1;
            zok = z > 0 || pz;
            z2 = z > 0 ? z - 1 : lz - 1;
        } else {
            yok = y + 1 < ly || py;
            y2 = y + 1 < ly ? y + 1 : 0;
            zok = z + 1 < lz || pz;
            z2 = z + 1 < lz ? z + 1 : 0;
        }
        const uint32_t rows[4] = {row, y2 + ly * z, y + ly * z2, y2 + ly * z2};
        const bool rok[4] = {true, yok, zok, yok && zok};
        const uint32_t nr = 1u + (yok ? 1u : 0u) + (zok ? 1u : 0u) + (yok && zok ? 1u : 0u);

        const uint32_t i0 = seg * kIsingSeg;
        const uint32_t nvalid = min(kIsingSeg, w - i0);
        const uint32_t off = row * pitch + i0;
        const u32x4 o4 = bload16(rw_own, off);
        uint32_t P[4] = {0u, 0u, 0u, 0u};  // per byte: up sites at position x over the 4 rows
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const u32x4 v = bload16(ro, rok[r] ? rows[r] * pitch + i0 : kNoRow);
            P[0] += v.x;
            P[1] += v.y;
            P[2] += v.z;
            P[3] += v.w;
        }
        // x-neighbour counts: position x-1 (colour 0) or x+1 (colour 1) of
        // the same rows.  gx: the one position whose x neighbour is absent.
        uint32_t xs[4];
        uint32_t gx = 0xFFFFFFFFu;
        if (colour == 0) {
            const bool xl_ok = i0 > 0 || px;
            const uint32_t xl = i0 > 0 ? i0 - 1 : w - 1;
            uint32_t e = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) e += bload8(ro, rok[r] && xl_ok ? rows[r] * pitch + xl : kNoRow);
            if (!xl_ok) gx = 0;
            xs[0] = __builtin_amdgcn_alignbyte(P[0], e << 24, 3);
            xs[1] = __builtin_amdgcn_alignbyte(P[1], P[0], 3);
            xs[2] = __builtin_amdgcn_alignbyte(P[2], P[1], 3);
            xs[3] = __builtin_amdgcn_alignbyte(P[3], P[2], 3);
        } else {
            const bool xr_in = i0 + kIsingSeg < w;
            const uint32_t xr = xr_in ? i0 + kIsingSeg : 0;  // position 0 across the end
            const bool xr_ok = xr_in || px;
            uint32_t e = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) e += bload8(ro, rok[r] && xr_ok ? rows[r] * pitch + xr : kNoRow);
            uint32_t m[4] = {P[0], P[1], P[2], P[3]};
            if (!xr_in) {
                const uint32_t pw = w - i0;  // 1..16 positions in this last segment
                if (!px) gx = pw - 1;
                if (pw != kIsingSeg) {
                    // byte pw (padding, zero) takes the wrapped neighbour of
                    // position w-1
#pragma unroll
                    for (uint32_t kk = 0; kk < 4; ++kk)
                        if ((pw >> 2) == kk) m[kk] |= e << (8u * (pw & 3u));
                    e = 0;
                }
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
        const U4 ra = philox(U4{i0 >> 4, row, (uint32_t)t, ctr_w(t, kIsingA, colour)}, k0, k1);
        const uint32_t rw[4] = {ra.x, ra.y, ra.z, ra.w};
        const uint32_t vbits = nvalid >= kIsingSeg ? 0xFFFFu : (1u << nvalid) - 1u;
        uint32_t valid[4], live[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) live[kk] = valid[kk] = nibble_bytes(vbits >> (4 * kk)) << 7;
        if (C.proposal) {
            const U4 rp =
                philox(U4{i0 >> 7, row, (uint32_t)t, ctr_w(t, kIsingProposal, colour)}, k0, k1);
            const uint32_t bits = (u4_word(rp, (i0 >> 5) & 3u) >> (i0 & 31u)) & 0xFFFFu;
            // Spin::rand proposes Up on bit 0 (state.rs:76-84); a proposal
            // equal to the current spin changes nothing.
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                live[kk] &= (~(nibble_bytes(bits >> (4 * kk)) ^ o[kk]) & kOne) << 7;
        }
        // absent slots: 8 - (present rows) * (present x offsets)
        const uint32_t a_in = 8u - 2u * nr, a_edge = 8u - nr;
        const uint32_t always0 = C.always[0], always1 = C.always[1];
        uint32_t T0[4], T1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            T0[k] = tab[a_in * 8u + k];
            T1[k] = tab[a_in * 8u + 4u + k];
        }
        uint32_t n[4], accH[4], tieH[4];
        decide16(cnt, o, rw, live, T0, T1, n, accH, tieH);
        // The position with an absent x neighbour (open x) has S = 2c + a_edge.
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
            if (tie_accept(lut, bj ? always1 : always0, bj, S, i0 + j, row, t, colour, k0, k1))
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
        if (colour == 1) {
            // bond sum over colour-1 sites: sum_j s_j (S_j - 8) with
            // s = 2n - 1 and S = 2c + a (a_edge at gx)
            uint32_t nc = 0, cs = 0;
            int32_t upc = 0;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint32_t cv = cnt[kk] & ((valid[kk] >> 7) * 0xFFu);
                nc = __builtin_amdgcn_udot4(n[kk], cv, nc, false);
                cs = __builtin_amdgcn_udot4(cv, kOne, cs, false);
                upc += __popc(n[kk]);
            }
            b_part = 4 * (int32_t)nc - 2 * (int32_t)cs +
                     ((int32_t)a_in - 8) * (2 * upc - (int32_t)nvalid);
            if (gx < nvalid) {
                const int32_t nj = (int32_t)((n[gx >> 2] >> (8u * (gx & 3u))) & 1u);
                b_part += (2 * nj - 1) * ((int32_t)a_edge - (int32_t)a_in);
            }
        }
        const int32_t up = __popc(n[0]) + __popc(n[1]) + __popc(n[2]) + __popc(n[3]);
        m_part = 2 * up - (int32_t)nvalid;
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

}  // namespace vg
