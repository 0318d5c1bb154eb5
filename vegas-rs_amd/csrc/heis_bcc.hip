// fp32 Heisenberg checkerboard colour pass on the body-centred cubic
// lattice, for any number of chains packed in one arena.
//
// The per-site update is the one of the sc pass (heis_common.hpp); only the
// neighbour sum differs.  A corner site of cell (x,y,z) (colour 0) bonds to
// the body centres of cells (x-dx, y-dy, z-dz), a body centre (colour 1) to
// the corners of cells (x+dx, y+dy, z+dz), dx,dy,dz in {0,1}; the field h is
// summed over the eight slots in the order dz, dy, dx (dx fastest) with an
// absent slot (open boundary) adding +0.0f.
//
// Work item: 4 consecutive positions of one row; each of the four neighbour
// rows is read with one aligned 16-byte load per component, plus the one
// position beyond the item (x-1 or x+4).  Same three phases as the sc pass
// (fields and first proposal calls, the wave's queue, updates).
#include "heis_common.hpp"

namespace vg {

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void heis_bcc_pass(const ChainDev* __restrict__ chains,
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
    const uint32_t ly = C.ly, lz = C.lz, w = C.w, pitch = C.pitch, pbc = C.pbc;
    const uint32_t row = fast_div(local, C.segs_m, C.segs_s);
    const uint32_t seg = local - row * C.segs;
    const uint32_t z = fast_div(row, C.ly_m, C.ly_s);
    const uint32_t y = row - z * ly;
    const bool px = pbc & 1u, py = pbc & 2u, pz = pbc & 4u;
    const uint32_t i0 = seg * kHeisSeg;
    const uint32_t nvalid = active ? min(kHeisSeg, w - i0) : 0u;
    const uint32_t own_off = active ? (row * pitch + i0) * 4u : kNoRow;
    uint32_t pend = 0;
    float os[4][3], hs[4][3], ps[4][3], us[4];
    {
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
        // slot rows in order (dy, dz) = (0,0), (1,0), (0,1), (1,1)
        const uint32_t rows[4] = {row, y2 + ly * z, y + ly * z2, y2 + ly * z2};
        const bool rok[4] = {true, yok, zok, yok && zok};
        // x offset of the second slot of each row: x-1 (colour 0), x+1 (colour 1)
        bool xe_ok;
        uint32_t xe_pos;
        if (colour == 0) {
            xe_pos = i0 > 0 ? i0 - 1 : w - 1;
            xe_ok = i0 > 0 || px;
        } else {
            xe_pos = i0 + kHeisSeg < w ? i0 + kHeisSeg : 0;
            xe_ok = i0 + kHeisSeg < w || px;
        }

        float4 so[3], v[4][3];
        // colour 1: the last segment's extra position is 0, the +x slot of
        // position w-1 across the periodic end (0 when x is open)
        float xe[4][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) so[c] = bload4f(rown[c], own_off);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t ro = rows[r] * pitch * 4u;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                v[r][c] = bload4f(roth[c], rok[r] ? ro + i0 * 4u : kNoRow);
                xe[r][c] = bload1f(roth[c], rok[r] && xe_ok ? ro + xe_pos * 4u : kNoRow);
            }
        }

        U4 fw = U4{0u, 0u, 0u, 0u};
        if (!C.proposal) fw = flip_words(C, i0, row, t, colour);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
#pragma unroll
            for (int c = 0; c < 3; ++c) os[j][c] = f4get(so[c], j);
            const uint32_t pos = i0 + j;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float acc_h = 0.0f;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float a = f4get(v[r][c], j);
                    float bx;
                    if (colour == 0)
                        bx = j == 0 ? xe[r][c] : f4get(v[r][c], j - 1);
                    else
                        bx = (pos + 1 == w || j == 3) ? xe[r][c] : f4get(v[r][c], j + 1);
                    acc_h = r == 0 ? fa(a, bx) : fa(fa(acc_h, a), bx);
                }
                hs[j][c] = acc_h;
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
            u32x4 vv;
            vv.x = __float_as_uint(ns[c][0]);
            vv.y = __float_as_uint(ns[c][1]);
            vv.z = __float_as_uint(ns[c][2]);
            vv.w = __float_as_uint(ns[c][3]);
            __builtin_amdgcn_raw_buffer_store_b128(vv, rown[c], own_off, 0, 0);
        }
    }

    heis_partials_sites(s32, bsum, partials, colour, 2u, s, total_blocks, b);
}

}  // namespace vg
