// Ising checkerboard colour pass on the simple-cubic lattice, multispin
// coded: 32 sites per 32-bit word, int8 or bit-packed planes, for any number
// of chains packed in one arena (a single chain is a batch of one).
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
// Work item: one 32-position segment of a row (bit j of every word below is
// position i0 + j), marched through K (the tile, a template parameter picked
// by the host) consecutive z planes, upwards in even chunks of K planes and
// downwards in odd ones.  Along z the row behind is the previous row's
// neighbour ahead, so each plane loads four rows (own, y-1, y+1, the next
// plane's) instead of six.  Int8 rows are packed to words on load (v_dot4 with bit
// weights) and expanded on store; bit-packed rows are one 32-bit load.  With
// bit-packed shadows of the two colour planes (int8 chains of a batch) every
// row is read as a shadow word, and the int8 row is only written.
//
// The decision works on whole words.  Each neighbour word is replaced by
// the complement of the site bits where that neighbour is absent (open
// boundary), and a carry-save adder counts, per site, the present neighbours
// aligned with it: a = 0..6 in three bit planes.  With g absent slots and
// p = 6 - g present, dE = 2J(2a - p) -+ 2h, so the class (b, a) with b the
// current bit fixes the threshold; the classes that are not always accepted
// come from the chain's plan for g (a few per temperature: a = 4, 5, 6 in a
// ferromagnet at h = 0).  For each such class the site mask is one bit
// operation on the count planes, and the draw's high byte is compared with
// the class threshold's high byte bit-serially: the eight high bits of the
// 32 sites' draws are the eight words of the two Philox calls keyed by
// positions i0 / 16 and i0 / 16 + 1 (bit j of word k is bit k of site j's
// high byte).  When a draw's high byte equals the threshold's (probability
// 2^-8) the low 24 bits decide; those ties leave a record (slot by ballot)
// in the wave's LDS queue and, once the wave has finished its planes, one
// lane per record draws them, stores the settled word and corrects the
// wave's sums.  The result equals comparing the full 32-bit draw.  No
// workgroup barrier is needed after the threshold table is loaded.
#include "ising_swar.hpp"

namespace vg {

constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kTieRows = 128;  // tie records per wave (one per row segment with ties)
constexpr uint32_t kTieWords = 6;   // meta, tie mask, count planes a0..a2, site bits

// 16 bytes of 0/1 -> 16 bits (byte j -> bit j).
__device__ __forceinline__ uint32_t pack_bytes16(u32x4 v) {
    const uint32_t lo = __builtin_amdgcn_udot4(
        v.y, 0x80402010u, __builtin_amdgcn_udot4(v.x, 0x08040201u, 0u, false), false);
    const uint32_t hi = __builtin_amdgcn_udot4(
        v.w, 0x80402010u, __builtin_amdgcn_udot4(v.z, 0x08040201u, 0u, false), false);
    return lo | (hi << 8);
}

// A row segment of 32 positions as loaded (int8: 32 bytes of 0/1;
// bit-packed: one word), and its site bits.  kNoRow reads 0.
template <bool BITS>
struct RowRaw {
    uint32_t w;
};
template <>
struct RowRaw<false> {
    u32x4 a, b;
};

// full = false: the segment's positions 16..31 are all padding (the last
// segment of a row of at most 16 more positions); their bytes are neither
// read (they read 0) nor written.
template <bool BITS>
__device__ __forceinline__ RowRaw<BITS> load_raw(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                 bool full = true) {
    RowRaw<BITS> v;
    if constexpr (BITS) {
        v.w = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    } else {
        v.a = bload16(r, off);
        v.b = bload16(r, full ? off + 16u : kNoRow);
    }
    return v;
}

template <bool BITS>
__device__ __forceinline__ uint32_t site_bits(const RowRaw<BITS>& v) {
    if constexpr (BITS)
        return v.w;
    else
        return pack_bytes16(v.a) | (pack_bytes16(v.b) << 16);
}

template <bool BITS>
__device__ __forceinline__ uint32_t load_word(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                              bool full = true) {
    return site_bits<BITS>(load_raw<BITS>(r, off, full));
}

template <bool BITS>
__device__ __forceinline__ void store_word(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t v,
                                           bool full = true) {
    if (BITS) {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0);
        return;
    }
    u32x4 a, b;
    a.x = nibble_bytes(v);
    a.y = nibble_bytes(v >> 4);
    a.z = nibble_bytes(v >> 8);
    a.w = nibble_bytes(v >> 12);
    b.x = nibble_bytes(v >> 16);
    b.y = nibble_bytes(v >> 20);
    b.z = nibble_bytes(v >> 24);
    b.w = nibble_bytes(v >> 28);
    __builtin_amdgcn_raw_buffer_store_b128(a, r, off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b128(b, r, full ? off + 16u : kNoRow, 0, 0);  // kNoRow: dropped
}

template <bool BITS>
__device__ __forceinline__ uint32_t word_off(uint32_t i0) {
    return BITS ? i0 >> 3 : i0;
}

// Bitwise function of three words given by its truth table: bit i of TT is
// the result for (a, b, c) = (i >> 2 & 1, i >> 1 & 1, i & 1).  One v_bitop3.
template <uint32_t TT>
__device__ __forceinline__ uint32_t bop3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, TT);
}
constexpr uint32_t kXor3 = 0x96;  // a ^ b ^ c
constexpr uint32_t kMaj3 = 0xE8;  // at least two of a, b, c

// One plan slot: of the sites in class mask e (restricted to the slot's
// bits), those whose draw high byte (planes d, bit k of site j = bit j of
// d[k]) is not below the slot's threshold byte are rejected, and those equal
// to it are ties when the threshold is nonzero.
__device__ __forceinline__ void plan_slot(uint32_t slot, uint32_t e, uint32_t o, uint32_t no,
                                          const uint32_t d[8], uint32_t& rej, uint32_t& tie) {
    const uint32_t cover = (slot >> 3) & 3u;
    const uint32_t m = cover == 3u ? e : e & (cover == 2u ? o : no);
    // d < t bit-serially from the least significant bit: at bit k,
    // lt = (~d & t) | (~(d ^ t) & lt)
    uint32_t lt = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) lt = bop3<0x8E>(d[k], 0u - ((slot >> (8 + k)) & 1u), lt);
    rej = bop3<0xF4>(rej, m, lt);  // rej | (m & ~lt)
    if (slot & 32u) {
        uint32_t eq = ~0u;  // eq & ~(d ^ t)
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) eq = bop3<0x90>(eq, d[k], 0u - ((slot >> (8 + k)) & 1u));
        tie = bop3<0xF8>(tie, m, eq);  // tie | (m & eq)
    }
}

// Sites whose aligned count (planes a0, a1, a2) is A: one v_bitop3.
template <uint32_t A>
__device__ __forceinline__ void plan_count(const uint32_t pw[14], uint32_t a0, uint32_t a1,
                                           uint32_t a2, uint32_t o, uint32_t no,
                                           const uint32_t d[8], uint32_t& rej, uint32_t& tie) {
    if (pw[2 * A] == 0u) return;
    const uint32_t e = bop3<(1u << A)>(a2, a1, a0);
    plan_slot(pw[2 * A], e, o, no, d, rej, tie);
    if (pw[2 * A + 1] != 0u) plan_slot(pw[2 * A + 1], e, o, no, d, rej, tie);
}

// Decision of the sites in `live` under the plan of absent count g: the
// accepted sites (ties excluded) and the tied ones.
__device__ __forceinline__ void decide(const ChainDev& C, uint32_t g, uint32_t a0, uint32_t a1,
                                       uint32_t a2, uint32_t o, uint32_t no, const uint32_t d[8],
                                       uint32_t live, uint32_t& acc, uint32_t& tie) {
    uint32_t pw[14];  // the plan's 14 words, one scalar load
#pragma unroll
    for (uint32_t k = 0; k < 14; ++k) pw[k] = C.plan[g][k >> 1][k & 1u];
    uint32_t rej = 0, t = 0;
    plan_count<0>(pw, a0, a1, a2, o, no, d, rej, t);
    plan_count<1>(pw, a0, a1, a2, o, no, d, rej, t);
    plan_count<2>(pw, a0, a1, a2, o, no, d, rej, t);
    plan_count<3>(pw, a0, a1, a2, o, no, d, rej, t);
    plan_count<4>(pw, a0, a1, a2, o, no, d, rej, t);
    plan_count<5>(pw, a0, a1, a2, o, no, d, rej, t);
    plan_count<6>(pw, a0, a1, a2, o, no, d, rej, t);
    acc = live & ~rej;
    tie = live & t;
}

// Settles the tied sites tm of one row segment: site j has bit b = bit j of
// ob, aligned count from a0..a2, absent slots g_row (+1 at position xe), so
// its class is S = 2c + g with c = a (b = 1) or 6 - g - a (b = 0).  Returns
// the accepted flips; dm, db collect their magnetisation and bond changes.
__device__ __forceinline__ uint32_t settle(const uint32_t* thr, const ChainDev& C, uint32_t tm,
                                           uint32_t a0, uint32_t a1, uint32_t a2, uint32_t ob,
                                           uint32_t g_row, uint32_t xe, uint32_t i0,
                                           uint32_t row, uint64_t t, uint32_t colour,
                                           uint32_t k0, uint32_t k1, int32_t& dm, int32_t& db) {
    uint32_t flips = 0;
    while (tm) {
        const uint32_t j = __builtin_ctz(tm);
        tm &= tm - 1u;
        const uint32_t bj = (ob >> j) & 1u;
        const uint32_t a = ((a0 >> j) & 1u) | (((a1 >> j) & 1u) << 1) | (((a2 >> j) & 1u) << 2);
        const uint32_t g = g_row + (j == xe ? 1u : 0u);
        const uint32_t S = 2u * (bj ? a : 6u - g - a) + g;
        if (tie_accept(thr, C.always[bj], bj, S, i0 + j, row, t, colour, k0, k1)) {
            flips |= 1u << j;
            const int32_t ds = 2 - 4 * (int32_t)bj;
            dm += ds;
            if (colour == 1) db += ds * ((int32_t)S - 6);  // bonds to colour 0
        }
    }
    return flips;
}

template <uint32_t K, uint32_t MODE>
__global__ __launch_bounds__(256) void ising_sc_pass(const ChainDev* __restrict__ chains,
                                                     const uint32_t* __restrict__ block_chain,
                                                     uint8_t* __restrict__ arena,
                                                     int32_t* __restrict__ partials,
                                                     uint32_t colour, uint32_t s,
                                                     uint32_t total_blocks, uint32_t grid,
                                                     uint32_t opts, uint32_t uniform_nb) {
    // storage of the own rows and of the neighbour rows; int8 planes with
    // shadows also store each updated row's bits in their own shadow
    constexpr bool BITS = MODE == kScBits;
    constexpr bool NB = MODE != kScInt8;
    constexpr bool SH = MODE == kScShadow;
    // per wave: kTieRows records of {meta, tie mask, a0, a1, a2, site bits}
    __shared__ uint32_t tie_q[kWaves][kTieRows][kTieWords];
    // opts: bit 0 = falling wave priorities; bits 8-15 = records the tie
    // queue may hold (at most kTieRows; fewer only to exercise the overflow
    // path in tests -- results do not depend on it).
    const uint32_t prio = opts & 1u;
    const uint32_t qcap = min((opts >> 8) & 0xFFu, kTieRows);

    // Blocks past the work (the grid is a multiple of 8) and chains that run
    // fewer sweeps in this block do nothing; they are not sent away here so
    // that every load below is issued at once.  Uniform over the workgroup.
    const uint32_t b = xcd_block(blockIdx.x, grid);
    const uint32_t bc = min(b, total_blocks - 1u);
    const ChainDev& C = chains[uniform_nb ? bc / uniform_nb : block_chain[bc]];
    // The descriptor's leading fields (all but the tables), loaded together:
    // a pass is short, and a chain of dependent loads at its start is a
    // good part of it.
    constexpr uint32_t kHead = offsetof(ChainDev, thr) / 4u;
    uint32_t hd[kHead];
#pragma unroll
    for (uint32_t k = 0; k < kHead; ++k) hd[k] = reinterpret_cast<const uint32_t*>(&C)[k];
#define VG_FIELD(f) hd[offsetof(ChainDev, f) / 4u]
#define VG_FIELD64(f) \
    (((uint64_t)hd[offsetof(ChainDev, f) / 4u + 1u] << 32) | hd[offsetof(ChainDev, f) / 4u])
    const bool active = b < total_blocks && s < VG_FIELD(limit);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t* const thr = &C.thr[0][0];  // read by tie settling only

    const uint64_t t = VG_FIELD64(sweep0) + s;
    const uint64_t seed = VG_FIELD64(seed);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t ly = VG_FIELD(ly), lz = VG_FIELD(lz), w = VG_FIELD(w);
    const uint32_t pitch = VG_FIELD(pitch), pbc = VG_FIELD(pbc);
    const uint32_t items = VG_FIELD(items), block0 = VG_FIELD(block0);
    const uint32_t segs = VG_FIELD(segs), segs_m = VG_FIELD(segs_m), segs_s = VG_FIELD(segs_s);
    const uint32_t ly_m = VG_FIELD(ly_m), ly_s = VG_FIELD(ly_s);
    const uint32_t plane = (uint32_t)VG_FIELD64(plane);
    const uint32_t spitch = VG_FIELD(spitch), rows = VG_FIELD(rows);
    const uint64_t soff_own = colour ? VG_FIELD64(soff[1]) : VG_FIELD64(soff[0]);
    const uint64_t soff_oth = colour ? VG_FIELD64(soff[0]) : VG_FIELD64(soff[1]);
    const uint64_t off_own = colour ? VG_FIELD64(off[1]) : VG_FIELD64(off[0]);
    const uint64_t off_oth = colour ? VG_FIELD64(off[0]) : VG_FIELD64(off[1]);
    const bool rnd = VG_FIELD(proposal) != 0;
#undef VG_FIELD
#undef VG_FIELD64
    // A colour plane is below 4 GiB (checked on the host): 32-bit offsets.
    uint8_t* const own_plane = arena + off_own;
    const uint8_t* const oth = arena + off_oth;
    const __amdgpu_buffer_rsrc_t rw_own = plane_rsrc(own_plane, plane);
    // neighbour rows: the other plane, or its shadow
    const __amdgpu_buffer_rsrc_t ro =
        SH ? plane_rsrc(arena + soff_oth, rows * spitch) : plane_rsrc(oth, plane);
    const __amdgpu_buffer_rsrc_t rw_sh = plane_rsrc(arena + soff_own, SH ? rows * spitch : 0u);
    const uint32_t zstep = ly * pitch;  // bytes between planes z and z+1 (own rows)
    const uint32_t npitch = SH ? spitch : pitch;  // the same for neighbour rows
    const uint32_t nzstep = ly * npitch;
    const uint32_t local = (b - block0) * kBlock + tid;
    const uint32_t wA = ctr_w(t, kIsingA, colour);
    uint32_t (*const q_wave)[kTieWords] = tie_q[wave];
    uint32_t nrec = 0;  // records queued by this wave, as counted by this lane

    int32_t m_part = 0, b_part = 0;
    uint32_t ups = 0;  // up sites after each plane's update, summed
    uint32_t seg = 0, y = 0, zc = 0;
    if (active && local < items) {
        // work items in natural order: segments of a row, rows y of a z
        // chunk, chunks; rows y and y+1 of one plane are neighbours in
        // memory (one 128-byte line for rows of 64 bytes) and in the grid
        const uint32_t r = fast_div(local, segs_m, segs_s);
        seg = local - r * segs;
        zc = fast_div(r, ly_m, ly_s);
        y = r - zc * ly;
    }
    if (active && local < items && y < ly) {
        const uint32_t z0 = zc * K;
        const uint32_t nk = min(K, lz - z0);
        const uint32_t i0 = seg * kScSeg;
        const uint32_t nvalid = min(kScSeg, w - i0);
        const uint32_t vmask = nvalid >= 32u ? 0xFFFFFFFFu : (1u << nvalid) - 1u;
        const bool px = pbc & 1u, py = pbc & 2u, pz = pbc & 4u;
        const bool ym_ok = y > 0 || py;
        const bool yp_ok = y + 1 < ly || py;
        // offsets of rows y-1 and y+1 relative to row y (wrapping when periodic)
        const int32_t dym = y > 0 ? -(int32_t)npitch : (int32_t)((ly - 1) * npitch);
        const int32_t dyp = y + 1 < ly ? (int32_t)npitch : -(int32_t)((ly - 1) * npitch);
        const uint32_t gy = (ym_ok ? 0u : 1u) + (yp_ok ? 0u : 1u);
        // x neighbour outside the 32 positions: i0-1 (q = 0) or i0+32 (q = 1)
        const bool xl_ok = i0 > 0 || px;
        const uint32_t xl_pos = i0 > 0 ? i0 - 1 : w - 1;
        const bool xr_in = i0 + kScSeg < w;
        // bit of the last valid position, which takes the right neighbour's bit
        const uint32_t xr_sh = xr_in ? 31u : nvalid - 1u;
        // Even z chunks march up, odd ones down: two chunks meeting at a seam
        // then read the seam planes in the same step, and the L2 fetches
        // those rows once.
        const bool down = zc & 1u;
        const uint32_t zfirst = down ? z0 + nk - 1u : z0;
        const int32_t dstep = down ? -(int32_t)zstep : (int32_t)zstep;
        const int32_t ndstep = down ? -(int32_t)nzstep : (int32_t)nzstep;
        uint32_t off = (y + ly * zfirst) * pitch + word_off<BITS>(i0);  // this segment's row
        uint32_t noff = (y + ly * zfirst) * npitch + word_off<NB>(i0);  // the same, neighbour rows
        // offset and presence of the row in plane z + dz (dz = +-1) of the
        // other colour, for the row at offset of in plane z
        const auto zrow = [&](uint32_t z, uint32_t of, bool up, bool& ok) {
            ok = up ? (z + 1 < lz || pz) : (z > 0 || pz);
            if (!ok) return kNoRow;
            return up ? (z + 1 < lz ? of + nzstep : of - (lz - 1) * nzstep)
                      : (z > 0 ? of - nzstep : of + (lz - 1) * nzstep);
        };
        // the row window along z: the plane behind the march and the current one
        bool behind_ok;
        const bool full = nvalid > 16u;  // the segment's upper 16 positions hold sites
        uint32_t behind = load_word<NB>(ro, zrow(zfirst, noff, down, behind_ok), full);
        uint32_t here = load_word<NB>(ro, noff, full);

        const uint32_t wP = ctr_w(t, kIsingProposal, colour);
        // Along the march: the plane ahead of plane z is z + 1 (up) or z - 1
        // (down), or across the periodic end from the lattice's last (up) or
        // first (down) plane, where it is absent when z is open.
        const uint32_t zend = down ? 0u : lz - 1u;
        const uint32_t nwrap = down ? (lz - 1) * nzstep : 0u - (lz - 1) * nzstep;
        const uint32_t rstep = down ? 0u - ly : ly;
        // x neighbours of position i: i-1, i (q = 0) or i, i+1 (q = 1) in the
        // other plane's same row; q alternates from plane to plane.  The bit
        // outside the segment is position i0-1 (q = 0), or i0+32, or position
        // 0 across the periodic end (q = 1); it reads 0 when x is open there,
        // and then xe is its site, the one with an absent x slot (32: none).
        const uint32_t xpos0 = xl_pos, xpos1 = xr_in ? i0 + kScSeg : 0u;
        const bool xok0 = xl_ok, xok1 = xr_in || px;
        const uint32_t xe0 = !xl_ok ? 0u : 32u, xe1 = !xr_in && !px ? nvalid - 1u : 32u;
        uint32_t z = zfirst, row = y + ly * zfirst;
        uint32_t q = (colour + y + zfirst) & 1u;
        // not unrolled: one plane's code is large, and the instruction
        // cache is shared by the waves of two CUs
#pragma unroll 1
        for (uint32_t k = 0; k < nk; ++k) {
            // The SIMD arbiter favours older waves; when every wave of the
            // pass is resident at once, a wave's priority falls as it
            // marches, so the waves of a SIMD progress together and the pass
            // does not end in a tail of a few young waves.
            if (prio) set_prio(((K - 1u - k) * 4u) / K);
            const bool at_end = z == zend;
            const bool ahead_ok = !at_end || pz;
            const uint32_t ahead_off = !at_end ? noff + ndstep : (pz ? noff + nwrap : kNoRow);
            // The own sites: their row as stored, or with shadows the own
            // colour's shadow word, which every write of the row keeps equal
            // to it (the int8 row is then only written here).
            uint32_t o_word = 0;
            if constexpr (SH)
                o_word = __builtin_amdgcn_raw_buffer_load_b32(rw_sh, noff, 0, 0);
            else
                o_word = site_bits<BITS>(load_raw<BITS>(rw_own, off, full));
            const RowRaw<NB> c_ym = load_raw<NB>(ro, ym_ok ? noff + dym : kNoRow, full);
            const RowRaw<NB> c_yp = load_raw<NB>(ro, yp_ok ? noff + dyp : kNoRow, full);
            const RowRaw<NB> c_ab = load_raw<NB>(ro, ahead_off, full);
            const uint32_t extra =
                load_pos<NB>(ro, noff - word_off<NB>(i0), q ? xpos1 : xpos0, q ? xok1 : xok0);

            // draws first: they do not wait for the loads.  Nothing branches
            // between the loads and the count planes, so the draws and the
            // loads' uses share one block and the wait falls after the draws.
            const U4 r0 = philox(U4{i0 >> 4, row, (uint32_t)t, wA}, k0, k1);
            const U4 r1 = philox(U4{(i0 >> 4) + 1u, row, (uint32_t)t, wA}, k0, k1);
            uint32_t d[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};

            const uint32_t o = o_word;
            const uint32_t ym = site_bits<NB>(c_ym);
            const uint32_t yp = site_bits<NB>(c_yp);
            const uint32_t ahead = site_bits<NB>(c_ab);
            const uint32_t g_row = gy + (behind_ok ? 0u : 1u) + (ahead_ok ? 0u : 1u);
            const uint32_t xe = q ? xe1 : xe0;
            const uint32_t eb = xe < 32u ? 1u << xe : 0u;  // site with an absent x neighbour
            const uint32_t no = ~o;
            const uint32_t mrow = here & vmask;
            // an absent neighbour counts as anti-aligned with its site
            const uint32_t xa = q ? mrow : ((((mrow << 1) | extra) & ~eb) | (no & eb));
            const uint32_t xb = q ? ((((mrow >> 1) | (extra << xr_sh)) & ~eb) | (no & eb)) : mrow;
            const uint32_t nym = ym_ok ? ym : no;
            const uint32_t nyp = yp_ok ? yp : no;
            const uint32_t nzm = behind_ok ? behind : no;
            const uint32_t nzp = ahead_ok ? ahead : no;
            // aligned count a = 0..6 of the present neighbours equal to the
            // site (n aligned where ~(n ^ o)): two full adders over aligned
            // triples, then their sum
            const uint32_t s1 = bop3<kXor3>(bop3<kXor3>(xa, xb, nym), o, ~0u);
            const uint32_t c1 = bop3<kXor3>(bop3<kMaj3>(xa, xb, nym), o, ~0u);
            const uint32_t s2 = bop3<kXor3>(bop3<kXor3>(nyp, nzm, nzp), o, ~0u);
            const uint32_t c2 = bop3<kXor3>(bop3<kMaj3>(nyp, nzm, nzp), o, ~0u);
            const uint32_t a0 = s1 ^ s2, h = s1 & s2;
            const uint32_t a1 = bop3<kXor3>(c1, c2, h);
            const uint32_t a2 = bop3<kMaj3>(c1, c2, h);
            // The compiler would otherwise sink the draws past the branch
            // below, behind the wait for the loads: an empty asm statement
            // taking them keeps them ahead of it.
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) asm volatile("" : "+v"(d[j]));
            uint32_t live = vmask;
            if (rnd) {
                // Spin::rand proposes Up when its draw is below 1/2
                // (state.rs:76-84): proposal bit 0 -> Up.  A proposal equal
                // to the current spin changes nothing.
                const U4 rp = philox(U4{i0 >> 7, row, (uint32_t)t, wP}, k0, k1);
                const uint32_t wsel = (i0 >> 5) & 3u;
                const uint32_t lo = wsel & 1u ? rp.y : rp.x, hi = wsel & 1u ? rp.w : rp.z;
                live &= ~((wsel & 2u ? hi : lo) ^ o);
            }

            // Sites of this row segment take the plan of their absent count:
            // g_row, and g_row + 1 at position xe.  Usually every lane of the
            // wave has one and the same count.
            uint32_t acc, tie;
            const uint32_t gmask = (1u << g_row) | (eb ? 2u << g_row : 0u);
            const uint32_t gu = __builtin_amdgcn_readfirstlane(g_row);
            if (__ballot(gmask != (1u << gu)) == 0) {
                decide(C, gu, a0, a1, a2, o, no, d, live, acc, tie);
            } else {
                acc = 0;
                tie = 0;
#pragma unroll 1
                for (uint32_t g = 0; g < kPlanG; ++g) {
                    if (__ballot((gmask >> g) & 1u) == 0) continue;
                    const uint32_t sel = (g_row == g ? ~eb : 0u) | (g_row + 1u == g ? eb : 0u);
                    uint32_t ag, tg;
                    decide(C, g, a0, a1, a2, o, no, d, live & sel, ag, tg);
                    acc |= ag;
                    tie |= tg;
                }
            }

            // Segments with ties leave one record each in the wave's queue
            // (their tied sites keep their bits in the store below); slots by
            // ballot.  When the queue is full, this lane settles them now.
            const uint64_t has = __ballot(tie != 0);
            if (has) {
                if (tie) {
                    const uint32_t slot =
                        nrec + __builtin_amdgcn_mbcnt_hi((uint32_t)(has >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)has, 0u));
                    if (slot < qcap) {
                        uint32_t* rec = q_wave[slot];
                        rec[0] = lane | ((z - z0) << 6) | (g_row << 9) | (xe << 12);
                        rec[1] = tie;
                        rec[2] = a0;
                        rec[3] = a1;
                        rec[4] = a2;
                        rec[5] = o ^ acc;
                    } else {
                        // Queue full (rare: a wave marching 8 planes near
                        // T_c queues a few dozen records against 128 slots).
                        int32_t dm = 0, db = 0;
                        acc |= settle(thr, C, tie, a0, a1, a2, o, g_row, xe, i0, row, t, colour,
                                      k0, k1, dm, db);
                    }
                }
                nrec += __builtin_popcountll(has);
            }
            const uint32_t n = o ^ acc;
            if (acc) {
                store_word<BITS>(rw_own, off, n, full);
                if (SH) __builtin_amdgcn_raw_buffer_store_b32(n, rw_sh, noff, 0, 0);
            }
            ups += __popc(n & vmask);
            if (colour == 1) {
                // sum over the segment's sites of s_j * (sum of its present
                // neighbours' spins) = 2a' - p, a' the count aligned with the
                // new bit: a for kept sites, p - a for flipped ones, p = 6 - g
                const uint32_t kept = vmask & ~acc;
                const int32_t sak = (int32_t)(__popc(a0 & kept) + 2u * __popc(a1 & kept) +
                                              4u * __popc(a2 & kept));
                const int32_t saf =
                    (int32_t)(__popc(a0 & acc) + 2u * __popc(a1 & acc) + 4u * __popc(a2 & acc));
                const int32_t nf = (int32_t)__popc(acc);
                const int32_t p = 6 - (int32_t)g_row;
                const int32_t xf = (acc & eb) ? 1 : 0, xk = (kept & eb) ? 1 : 0;
                b_part += 2 * (sak - saf) - p * ((int32_t)nvalid - 2 * nf) + xk - xf;
            }
            behind = here;
            here = ahead;
            behind_ok = true;
            off += dstep;
            noff += ndstep;
            z = down ? z - 1u : z + 1u;
            row += rstep;
            q ^= 1u;
        }
        m_part += 2 * (int32_t)ups - (int32_t)(nk * nvalid);
    }

    // Settle this wave's queued ties: one lane per record draws the low 24
    // bits of each tied site, stores the settled word and corrects the sums.
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
            const uint32_t owner = (wave << 6) | (meta & 63u), kz = (meta >> 6) & 7u;
            const uint32_t g_row = (meta >> 9) & 7u, xe = (meta >> 12) & 63u;
            const uint32_t l = (b - block0) * kBlock + owner;
            const uint32_t r = fast_div(l, segs_m, segs_s);
            const uint32_t sg = l - r * segs;
            const uint32_t zq = fast_div(r, ly_m, ly_s);
            const uint32_t yy = r - zq * ly;
            const uint32_t zz = zq * K + kz;
            const uint32_t row = yy + ly * zz;
            const uint32_t i0 = sg * kScSeg;
            const uint32_t bits = rec[5];
            const uint32_t flips = settle(thr, C, rec[1], rec[2], rec[3], rec[4], bits, g_row, xe,
                                          i0, row, t, colour, k0, k1, m_part, b_part);
            if (flips) {
                // the owner stored the segment with the tied sites unchanged
                if (SH)
                    __builtin_amdgcn_raw_buffer_store_b32(bits ^ flips, rw_sh,
                                                          row * spitch + (i0 >> 3), 0, 0);
                if (BITS) {
                    __builtin_amdgcn_raw_buffer_store_b32(bits ^ flips, rw_own,
                                                          row * pitch + word_off<BITS>(i0), 0, 0);
                } else {
                    uint32_t f = flips;
                    while (f) {
                        const uint32_t j = __builtin_ctz(f);
                        f &= f - 1u;
                        own_plane[row * pitch + i0 + j] = (uint8_t)(((bits >> j) & 1u) ^ 1u);
                    }
                }
            }
        }
    }

    m_part = wave_sum(m_part);
    b_part = wave_sum(b_part);
    if (active && lane == 0) {
        int32_t* p = partials + ((uint64_t)(s * 2u + colour) * total_blocks + b) * (2u * kWaves) +
                     2u * wave;
        p[0] = m_part;
        p[1] = b_part;
    }
}

template __global__ void ising_sc_pass<1, kScInt8>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<2, kScInt8>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<4, kScInt8>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<8, kScInt8>(const ChainDev*, const uint32_t*, uint8_t*,
                                                  int32_t*, uint32_t, uint32_t, uint32_t,
                                                  uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<1, kScBits>(const ChainDev*, const uint32_t*, uint8_t*,
                                                 int32_t*, uint32_t, uint32_t, uint32_t,
                                                 uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<2, kScBits>(const ChainDev*, const uint32_t*, uint8_t*,
                                                 int32_t*, uint32_t, uint32_t, uint32_t,
                                                 uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<4, kScBits>(const ChainDev*, const uint32_t*, uint8_t*,
                                                 int32_t*, uint32_t, uint32_t, uint32_t,
                                                 uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<8, kScBits>(const ChainDev*, const uint32_t*, uint8_t*,
                                                 int32_t*, uint32_t, uint32_t, uint32_t,
                                                 uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<1, kScShadow>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<2, kScShadow>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<4, kScShadow>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t, uint32_t);
template __global__ void ising_sc_pass<8, kScShadow>(const ChainDev*, const uint32_t*, uint8_t*,
                                                     int32_t*, uint32_t, uint32_t, uint32_t,
                                                     uint32_t, uint32_t, uint32_t);

// Packs the int8 planes of an sc chain into its shadows: one thread per
// shadow word (bit j of word k of a row = position 32k + j; positions past
// the row read 0).
__global__ void ising_shadow_build(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint32_t wpr = C.spitch / 4u;  // words per shadow row
    const uint64_t per = (uint64_t)C.rows * wpr;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2u * per) return;
    const uint32_t c = (uint32_t)(idx / per);
    const uint64_t e = idx - c * per;
    const uint32_t row = (uint32_t)(e / wpr), k = (uint32_t)(e % wpr);
    const uint8_t* src = arena + C.off[c] + (uint64_t)row * C.pitch;
    uint32_t v = 0;
    for (uint32_t j = 0; j < 32u; ++j) {
        const uint32_t i = 32u * k + j;
        if (i < C.w && src[i]) v |= 1u << j;
    }
    *reinterpret_cast<uint32_t*>(arena + C.soff[c] + (uint64_t)row * C.spitch + 4u * k) = v;
}

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
