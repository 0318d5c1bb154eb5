// Byte-wise (four sites per 32-bit register) pieces of the int8 Ising
// decision shared by the colour passes: comparisons of a draw's high bytes
// with threshold high bytes, the class-table lookup by v_perm_b32, and the
// tie rule.
//
// A site's flip is accepted when its class (current bit b, S = 2 * (up
// neighbours) + (absent neighbour slots)) is always accepted, or when its
// 32-bit draw is below ceil(exp(-dE/T) * 2^32) (the reference's
// `u < exp(-dE/T)`, src/integrator.rs:127-134, with u quantised to 2^-32).
// The high byte of the draw is compared first; only on equal high bytes
// (probability 2^-8) do the low 24 bits decide.  An always-accepted class
// has table byte 0xFF: a draw byte below it accepts, and 0xFF reads as a tie
// that the tie rule accepts without drawing.
#pragma once

#include "vg_device.hpp"

namespace vg {

constexpr uint32_t kH = 0x80808080u;  // byte MSBs
constexpr uint32_t kL = 0x7F7F7F7Fu;
constexpr uint32_t kOne = 0x01010101u;

// Byte-wise unsigned r < t, as MSB-per-byte flags.
__device__ __forceinline__ uint32_t bytes_lt(uint32_t r, uint32_t t) {
    const uint32_t d = (r | kH) - (t & kL);
    return ((~r & t) | (~(r ^ t) & ~d)) & kH;
}

// Byte-wise r == t, as MSB-per-byte flags.
__device__ __forceinline__ uint32_t bytes_eq(uint32_t r, uint32_t t) {
    const uint32_t x = r ^ t;
    return ~(((x & kL) + kL) | x) & kH;
}

// Four bits -> one byte each (0 or 1).
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t v) {
    return ((v & 0xFu) * 0x00204081u) & kOne;
}

// Byte MSB flags of 16 sites -> 16-bit mask (byte j -> bit j).
__device__ __forceinline__ uint32_t msb_mask16(const uint32_t h[4]) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) m |= ((((h[k] >> 7) & kOne) * 0x10204080u) >> 28) << (4 * k);
    return m;
}

// Low 24 bits of a tied draw decide: accept iff they are below the low 24
// bits of the threshold; an always-accepted class needs no draw.  The draw
// of position pos is word pos & 3 of the call (pos >> 2, row, sweep,
// purpose IsingB).
__device__ __forceinline__ bool tie_accept(const uint32_t* lut, uint32_t always_b, uint32_t bj,
                                           uint32_t S, uint32_t pos, uint32_t row, uint64_t t,
                                           uint32_t colour, uint32_t k0, uint32_t k1) {
    if ((always_b >> S) & 1u) return true;
    const U4 rb = philox(U4{pos >> 2, row, (uint32_t)t, ctr_w(t, kIsingB, colour)}, k0, k1);
    return (u4_word(rb, pos & 3u) >> 8) < (lut[bj * kThrS + S] & 0xFFFFFFu);
}

// The two storage forms of a colour plane.  Int8: one byte per position,
// row segments of 16 positions are 16-byte loads.  Bit-packed: bit (i & 7)
// of byte i >> 3 of the row, a segment is one 16-bit load expanded to the
// same 16 bytes in registers, so everything after the loads is shared.
template <bool BITS>
__device__ __forceinline__ uint32_t seg_off(uint32_t i) {
    return BITS ? i >> 3 : i;
}

template <bool BITS>
__device__ __forceinline__ u32x4 load_seg(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if (!BITS) return bload16(r, off);
    const uint32_t v = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    u32x4 e;
    e.x = nibble_bytes(v);
    e.y = nibble_bytes(v >> 4);
    e.z = nibble_bytes(v >> 8);
    e.w = nibble_bytes(v >> 12);
    return e;
}

// Spin bit of position i of the row starting at byte offset roff (kNoRow:
// absent, reads 0).
template <bool BITS>
__device__ __forceinline__ uint32_t load_pos(__amdgpu_buffer_rsrc_t r, uint32_t roff, uint32_t i,
                                             bool present) {
    const uint32_t v = bload8(r, present ? roff + seg_off<BITS>(i) : kNoRow);
    return BITS ? (v >> (i & 7u)) & 1u : v;
}

// 16 bytes of 0/1 -> 16 bits (byte j -> bit j).
__device__ __forceinline__ uint32_t pack16(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return (((a & kOne) * 0x10204080u) >> 28) | ((((b & kOne) * 0x10204080u) >> 28) << 4) |
           ((((c & kOne) * 0x10204080u) >> 28) << 8) | ((((d & kOne) * 0x10204080u) >> 28) << 12);
}

// Class tables of the bcc and fcc passes, built in LDS per workgroup: for
// every absent-slot count a = 0..z and bit b, the 16 high bytes (4 words)
// indexed by the up-neighbour count c, class S = 2c + a (0 where S > 2z).
// tab must hold (z + 1) * 8 words; every thread of the workgroup calls this
// before the barrier that publishes lut.
__device__ __forceinline__ void class_tables(const ChainDev& C, uint32_t z, uint32_t* tab) {
    const uint32_t tid = threadIdx.x;
    if (tid < (z + 1u) * 8u) {
        const uint32_t a = tid >> 3, b = (tid >> 2) & 1u, wd = tid & 3u;
        uint32_t v = 0;
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            const uint32_t S = 2u * (4u * wd + e) + a;
            uint32_t byte = 0;
            if (S <= 2u * z)
                byte = ((C.always[b] >> S) & 1u) ? 0xFFu : C.thr[b][S] >> 24;
            v |= byte << (8u * e);
        }
        tab[tid] = v;
    }
}

// Decision of 16 sites from their up-neighbour counts cnt (bytes 0..15),
// current bits o, draw words rw (high bytes) and live flags (byte MSBs: a
// valid site whose proposal changes it), with the tables T0 / T1 (4 words,
// for b = 0 / 1) of the item's absent count.  Returns the new bits with the
// accepted flips applied (tied sites unchanged) and, in acc / tie, the MSB
// flags of the accepted and tied sites.
__device__ __forceinline__ void decide16(const uint32_t cnt[4], const uint32_t o[4],
                                         const uint32_t rw[4], const uint32_t live[4],
                                         const uint32_t T0[4], const uint32_t T1[4],
                                         uint32_t n[4], uint32_t acc[4], uint32_t tie[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t lo = cnt[k] & 0x07070707u;  // entry within a half table
        const uint32_t hi = __builtin_amdgcn_perm(0u, 0xFF00u, (cnt[k] >> 3) & kOne);  // c >= 8
        const uint32_t a0 = __builtin_amdgcn_perm(T0[1], T0[0], lo);
        const uint32_t b0 = __builtin_amdgcn_perm(T0[3], T0[2], lo);
        const uint32_t a1 = __builtin_amdgcn_perm(T1[1], T1[0], lo);
        const uint32_t b1 = __builtin_amdgcn_perm(T1[3], T1[2], lo);
        const uint32_t th0 = (b0 & hi) | (a0 & ~hi);
        const uint32_t th1 = (b1 & hi) | (a1 & ~hi);
        const uint32_t bm = __builtin_amdgcn_perm(0u, 0xFF00u, o[k]);  // 0xFF where Up
        const uint32_t th = (th1 & bm) | (th0 & ~bm);
        acc[k] = bytes_lt(rw[k], th) & live[k];
        tie[k] = bytes_eq(rw[k], th) & live[k];
        n[k] = o[k] ^ (acc[k] >> 7);
    }
}

}  // namespace vg
