// fp32 Heisenberg colour pass on the face-centred cubic lattice (four colour
// passes per sweep, one per sublattice; geometry in fcc_common.hpp), for any
// number of chains packed in one arena.
//
// The per-site update is the one of the sc and bcc passes (heis_common.hpp);
// the field h sums the twelve neighbour slots in slot order, v0 + v1 first,
// an absent slot (open boundary) adding +0.0f, and the bond sum of the pass
// of colour a uses the same sum over the slots of the lower colours b < a
// (the first 4a slots), so over a sweep every bond counts once.
//
// Work item: 4 consecutive positions of one row.  Each other sublattice b is
// read in turn: four neighbour rows at the same x (a and b differ along y and
// z), or two rows at x and at the x-offset position (they differ along x),
// one aligned 16-byte load per component and row plus the one position
// beyond the item.  Same three phases as the sc pass (fields and first
// proposal calls, the wave's queue, updates).
#include "fcc_common.hpp"
#include "heis_common.hpp"

namespace vg {

template <uint32_t A>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void heis_fcc_pass(const ChainDev* __restrict__ chains,
                                                     const uint32_t* __restrict__ block_chain,
                                                     float* __restrict__ arena,
                                                     int64_t* __restrict__ partials, uint32_t s,
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
    // Sublattice planes through buffer resources: 32-bit byte offsets (a
    // component plane is below 4 GiB, checked on the host), and kNoRow reads
    // 0 for an absent neighbour row.  Every lane runs the same code: a lane
    // past the chain's items (the tail of its last block) has no valid site
    // and its own row offset is kNoRow, so it reads zeros and its stores are
    // dropped; a site j >= nvalid of a short last segment is padding, which
    // every kernel keeps at zero.  Neither is ever accepted, so both add
    // zero to the sums.
    const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
    const uint64_t plane = C.plane;
    const uint32_t row = fast_div(local, C.segs_m, C.segs_s);
    const uint32_t seg = local - row * C.segs;
    const uint32_t z = fast_div(row, C.ly_m, C.ly_s);
    const uint32_t y = row - z * ly;
    const bool px = pbc & 1u, py = pbc & 2u, pz = pbc & 4u;
    const uint32_t i0 = seg * kHeisSeg;
    const uint32_t nvalid = active ? min(kHeisSeg, w - i0) : 0u;
    const uint32_t own_off = active ? (row * pitch + i0) * 4u : kNoRow;
    const auto rsrc = [&](uint32_t sub, int c) {
        return plane_rsrc(reinterpret_cast<const uint8_t*>(arena + C.off[sub] + c * plane),
                          (uint32_t)(plane * 4u));
    };
    __amdgpu_buffer_rsrc_t rown[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) rown[c] = rsrc(A, c);
    uint32_t pend = 0;
    float os[4][3], hs[4][3], hb[4][3], ps[4][3], us[4];
    {
        // x-offset position beyond the item: i0-1 (colours 0, 1) or i0+4,
        // or position 0 across the periodic end (colours 2, 3: the +x slot
        // of position w-1, whichever j it has; 0 when x is open)
        constexpr bool kXneg = A < 2;
        const bool xr_in = i0 + kHeisSeg < w;
        const uint32_t xe_pos = kXneg ? (i0 > 0 ? i0 - 1 : w - 1) : (xr_in ? i0 + kHeisSeg : 0);
        const bool xe_ok = kXneg ? (i0 > 0 || px) : (xr_in || px);

        float4 so[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) so[c] = bload4f(rown[c], own_off);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) os[j][c] = f4get(so[c], j);

        uint32_t k = 0;  // slots summed so far
#pragma unroll
        for (uint32_t ob = 0; ob < 4; ++ob) {
            if (ob == A) continue;
            const uint32_t diff = fcc_diff(A, ob);
            __amdgpu_buffer_rsrc_t roth[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) roth[c] = rsrc(ob, c);
            bool yok, zok;
            const uint32_t y2 = fcc_move(y, fcc_step(A, ob, 1), ly, py, yok);
            const uint32_t z2 = fcc_move(z, fcc_step(A, ob, 2), lz, pz, zok);
            if (!(diff & 1u)) {
                // y and z differ: slots (y,z), (y',z), (y,z'), (y',z') at x
                const uint32_t rows[4] = {row, y2 + ly * z, y + ly * z2, y2 + ly * z2};
                const bool rok[4] = {true, yok, zok, yok && zok};
                float4 v[4][3];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        v[r][c] = bload4f(roth[c], rok[r] ? (rows[r] * pitch + i0) * 4u : kNoRow);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const float vv = f4get(v[r][c], j);
                            hs[j][c] = k == 0 ? vv : fa(hs[j][c], vv);
                        }
                    ++k;
                }
            } else {
                // x and one of y, z differ: slots (x, R0), (x', R0), (x, R1),
                // (x', R1) with R1 the row shifted along y or z
                const bool ok2 = (diff & 2u) ? yok : zok;
                const uint32_t row2 = (diff & 2u) ? y2 + ly * z : y + ly * z2;
                const uint32_t rows[2] = {row, row2};
                const bool rok[2] = {true, ok2};
                float4 v[2][3];
                float xe[2][3];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const uint32_t ro = rows[r] * pitch * 4u;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        v[r][c] = bload4f(roth[c], rok[r] ? ro + i0 * 4u : kNoRow);
                        xe[r][c] = bload1f(roth[c], rok[r] && xe_ok ? ro + xe_pos * 4u : kNoRow);
                    }
                }
#pragma unroll
                for (int r = 0; r < 2; ++r) {
#pragma unroll
                    for (int d1 = 0; d1 < 2; ++d1) {
#pragma unroll
                        for (uint32_t j = 0; j < 4; ++j) {
                            const uint32_t pos = i0 + j;
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                float vv;
                                if (d1 == 0)
                                    vv = f4get(v[r][c], j);
                                else if (kXneg)
                                    vv = j == 0 ? xe[r][c] : f4get(v[r][c], j - 1);
                                else
                                    vv = (pos + 1 == w || j == 3) ? xe[r][c] : f4get(v[r][c], j + 1);
                                hs[j][c] = k == 0 ? vv : fa(hs[j][c], vv);
                            }
                        }
                        ++k;
                    }
                }
            }
            if (A > 0 && ob + 1 == A) {
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j)
#pragma unroll
                    for (int c = 0; c < 3; ++c) hb[j][c] = hs[j][c];
            }
        }
        if (A == 0) {  // no lower colour: the bond share is not summed
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
#pragma unroll
                for (int c = 0; c < 3; ++c) hb[j][c] = 0.0f;
        }

        U4 fw = U4{0u, 0u, 0u, 0u};
        if (!C.proposal) fw = flip_words(C, i0, row, t, A);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
            heis_propose(C, A, i0 + j, row, t, fw, j, os[j], us[j], ps[j], pend);
        pend &= (1u << nvalid) - 1u;  // no further calls for sites that do not exist
    }

    // Phase 2: the wave's sites whose first proposal call missed.
    if (C.proposal) heis_resolve(C, hq[threadIdx.x >> 6], pend, i0, row, t, A, ps);

    // Phase 3: accept or reject, store, sums.
    {
        float ns[3][4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            float n[3];
            heis_update(C, A > 0, os[j], hs[j], hb[j], us[j], ps[j], n, s32, bsum, j < nvalid);
            ns[0][j] = n[0];
            ns[1][j] = n[1];
            ns[2][j] = n[2];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u32x4 vv;
            vv.x = __float_as_uint(ns[c][0]);
            vv.y = __float_as_uint(ns[c][1]);
            vv.z = __float_as_uint(ns[c][2]);
            vv.w = __float_as_uint(ns[c][3]);
            __builtin_amdgcn_raw_buffer_store_b128(vv, rown[c], own_off, 0, 0);
        }
    }

    heis_partials_sites(s32, bsum, partials, A, 4u, s, total_blocks, b);
}

template __global__ void heis_fcc_pass<0>(const ChainDev*, const uint32_t*, float*, int64_t*,
                                          uint32_t, uint32_t, uint32_t);
template __global__ void heis_fcc_pass<1>(const ChainDev*, const uint32_t*, float*, int64_t*,
                                          uint32_t, uint32_t, uint32_t);
template __global__ void heis_fcc_pass<2>(const ChainDev*, const uint32_t*, float*, int64_t*,
                                          uint32_t, uint32_t, uint32_t);
template __global__ void heis_fcc_pass<3>(const ChainDev*, const uint32_t*, float*, int64_t*,
                                          uint32_t, uint32_t, uint32_t);

}  // namespace vg
