// Shared device-side definitions of the MI355X Metropolis engine: the
// Philox4x32-10 generator, the counter layout that keys every random draw by
// (chain seed, sweep, colour, row, position), the chain descriptor the
// kernels read, and the colour-split spin layout.
//
// Layout of one chain (sc lattice lx*ly*lz, bcc and fcc lattices lx*ly*lz
// cells):
//   The colours are independent sets of the lattice: sc and bcc are
//   bipartite (two colours), fcc takes four.  Colour c of sc site (x,y,z) is
//   (x+y+z)&1; of a bcc or fcc site it is its sublattice.  Each colour is
//   stored as its own plane of rows; row r = y + ly*z; within a row,
//   position i holds
//     sc     : the colour-c site at x = 2*i + ((c+y+z)&1)   (w = lx/2 positions)
//     bcc/fcc: the sublattice-c site of cell x = i           (w = lx positions)
//   Rows are `pitch` elements apart (pitch is a multiple of 64 bytes), so a
//   row is read with aligned 16-byte loads.  Ising planes hold one byte per
//   site (1 = Up, 0 = Down); Heisenberg chains hold three fp32 planes (x, y,
//   z components) per colour.  Padding elements past w stay zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VG_HD __host__ __device__ __forceinline__

namespace vg {

// ---------------------------------------------------------------------------
// Philox4x32-10, Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11.  Multipliers and Weyl key increments as published
// there.  counter (x,y,z,w), key (k0,k1) -> 4 x 32 random bits.
// ---------------------------------------------------------------------------
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct U4 {
    uint32_t x, y, z, w;
};

// Three-way xor: one v_bitop3_b32 (truth table 0x96) in device code.
VG_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

VG_HD U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        U4 n;
        n.x = xor3((uint32_t)(p1 >> 32), c.y, k0);
        n.y = (uint32_t)p1;
        n.z = xor3((uint32_t)(p0 >> 32), c.w, k1);
        n.w = (uint32_t)p0;
        c = n;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

VG_HD uint32_t u4_word(const U4& v, uint32_t k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// Counter word w: bit 0 colour & 1, bits 1-6 purpose, bit 7 colour >> 1
// (fcc colours 2 and 3), bits 8-31 sweep >> 32.
// Counter words x, y, z: position group, row, sweep & 0xffffffff (or call
// index for the initial state, which has no sweep).
enum Purpose : uint32_t {
    kIsingA = 0,         // high 8 bits of the acceptance draw, 16 positions per call
    kIsingB = 1,         // low 24 bits of the acceptance draw, 4 positions per call
    kIsingProposal = 2,  // random-proposal bit, 128 positions per call
    kIsingInit = 3,      // initial spin, 4 positions per call
    kHeisInit = 4,       // initial spin stream, one position per stream
    kWolffBond = 5,      // Wolff: a bond's activation draw, keyed by its lower site and slot
    kWolffSeed = 6,      // Wolff: the cluster's seed site, one call per step
    kGraphIsing = 8,     // graph chains, Ising sweep: acceptance (word 0), proposal (word 1)
    kGraphInit = 9,      // graph chains, Ising initial spin
    kHeisStream = 16,    // Heisenberg sweep: proposal call k of a site uses purpose 16 + k
};
constexpr uint32_t kHeisMaxCalls = 16;
constexpr uint32_t kHeisInitMaxCalls = 64;

VG_HD uint32_t ctr_w(uint64_t sweep, uint32_t purpose, uint32_t colour) {
    return ((uint32_t)(sweep >> 32) << 8) | ((colour >> 1) << 7) | (purpose << 1) | (colour & 1u);
}

// ---------------------------------------------------------------------------
// Fixed point used for the Heisenberg per-sweep sums: every per-site fp32
// contribution (|v| <= 8) is rounded to an int32 multiple of 2^-27 and summed
// in int64, so the sums do not depend on reduction order.
// ---------------------------------------------------------------------------
constexpr float kFixScale = 134217728.0f;  // 2^27

// Ising acceptance classes per current bit: S = 0 .. 2z, z <= 12.
constexpr uint32_t kThrS = 25;
// Class plans of the multispin sc pass: absent-slot counts 0..5 of a site.
constexpr uint32_t kPlanG = 6;

// ---------------------------------------------------------------------------
// Chain descriptor as the kernels see it (one per chain of a kernel group).
// ---------------------------------------------------------------------------
struct ChainDev {
    uint64_t seed;         // Philox key (k0 = low word, k1 = high word)
    uint64_t sweep0;       // sweep counter at the start of the current block
    uint64_t off[4];       // element offset of colour plane c in the arena (ncol planes)
    uint64_t plane;        // elements per plane (rows * pitch); Heisenberg y/z planes follow x
    uint32_t lx, ly, lz, kind;
    uint32_t w, pitch, rows, pbc;
    uint32_t segs;         // work items per row
    uint32_t items;        // work items per pass (rows * segs)
    uint32_t block0;       // first block of this chain in the pass grid
    uint32_t nblocks;      // blocks of this chain
    uint32_t proposal;     // 0 flip, 1 random
    uint32_t slot;         // chain index in its batch (record row)
    uint32_t z;            // coordination number: 6 (sc), 8 (bcc) or 12 (fcc)
    uint32_t ncol;         // colours (planes): 2 (sc, bcc) or 4 (fcc)
    uint32_t limit;        // sweeps of the current block this chain runs (the rest skip it)
    // n / d == (n * m) >> s for n < 2^31 (segs, ly and hy divisors)
    uint32_t segs_m, segs_s, ly_m, ly_s;
    // Ising work order (bcc, fcc): rows of one y parity, then the other, a
    // tile of z planes per lane (the sc pass takes the rows in natural order)
    uint32_t hy;           // (ly + 1) / 2, rows of one parity class per plane
    uint32_t hy_m, hy_s;
    uint32_t nzc;          // z chunks of one tile of planes
    // int8 sc chains of a batch: bit-packed shadows of the two colour planes
    // (bit j of word k of a row = position 32k + j; rows of spitch bytes at
    // byte offsets soff[c] of the arena), equal to the int8 planes after
    // every kernel that writes those; the colour passes read the neighbour
    // rows from the other colour's shadow.  shadow = 0: no shadows.
    uint32_t spitch, shadow;
    uint64_t soff[2];
    // Ising acceptance: class (b, S) with b the current bit and S the sum over
    // neighbour slots of 2*b_j (present) or 1 (absent), S in [0, 2z].
    uint32_t thr[2][kThrS];  // accept iff draw < thr (32-bit draw)
    uint32_t always[2];    // bit S set: class always accepted
    // The multispin sc decision's class plans, per absent-slot count g =
    // 0..5 of a site (p = 6 - g present neighbours) and aligned count a =
    // 0..p (present neighbours equal to the site's bit b): up to two slots
    // for the classes (b, a) that are not always accepted, 0 for none.  Slot
    // bits: 3 covers b = 0, 4 covers b = 1, 5 the threshold is nonzero (a
    // draw whose high byte equals bits 8-15 is a tie), 8-15 the threshold's
    // high byte.  One slot covers both bits when those fields agree.
    uint32_t plan[kPlanG][7][2];
    // Heisenberg parameters, fp32.
    float jn;              // -J
    float kk;              // K
    float hm;              // |h|
    float beta;            // (float)(1/T)
    float ax, ay, az;      // anisotropy axis
    float fx, fy, fz;      // field orientation
    float pad1, pad2;
    // Wolff chains (natural site order): sites, and the bond threshold --
    // a bond is active iff its 32-bit draw < wolff_thr, or always.
    uint64_t nsites;
    uint32_t wolff_thr, wolff_all;
    uint32_t lx_m, lx_s;   // n / lx == (n * lx_m) >> lx_s for n < 2^31 (Wolff cells)
    // Graph chains (byte offsets in the arena): off[0] spins (Ising 1 byte
    // per position; Heisenberg x, y, z planes of nsites floats), off[1]
    // merged-row offsets (u32, nsites + 1), off[2] neighbour positions
    // (u32), off[3] couplings a (f64 Ising, f32 Heisenberg), goff[0]
    // lower-colour couplings w, goff[1] site id of each position (u32),
    // goff[2] colour class starts (u32, ncol + 1).  Ising parameters in f64:
    // 1/T, |h| (0 without Zeeman) and the field's sign.
    uint64_t goff[3];
    double gbeta, ghm, ghz;
};

constexpr uint32_t kBlock = 256;       // threads per workgroup of a pass
constexpr uint32_t kIsingSeg = 16;     // positions per Ising row segment (bcc, fcc)
constexpr uint32_t kScSeg = 32;        // positions per sc Ising work item (one word of site bits)
// sc pass modes: int8 planes; bit-packed planes; int8 planes whose
// neighbour rows are read from bit-packed shadows.
enum ScMode : uint32_t { kScInt8 = 0, kScBits = 1, kScShadow = 2 };
constexpr uint32_t kHeisSeg = 4;       // positions per Heisenberg work item
constexpr uint32_t kNumXcd = 8;        // MI355X: 8 XCDs, workgroups dealt round-robin
// Wolff chains of at most this many sites keep their spins and visit marks
// in LDS, one bit per site each (2 x 64 KiB of the 160 KiB of a CU), and the
// first kWolffLdsQueue entries of the cluster queue (31 KiB).
constexpr uint32_t kWolffLdsSites = 1u << 19;
constexpr uint32_t kWolffLdsQueue = 7936;
constexpr uint32_t kWolffLdsWordsMax = 2u * (kWolffLdsSites / 32u) + kWolffLdsQueue;

// Exact unsigned division by a per-chain constant d for n < 2^31:
// m = ceil(2^s / d) with s = 31 + ceil(log2 d) (computed on the host).
__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m, uint32_t s) {
    return (uint32_t)(((uint64_t)n * m) >> s);
}

// Workgroups are dispatched round-robin over the 8 XCDs; map hardware block
// b to a logical block so that each XCD walks one contiguous slab of the work
// (its L2 then holds the rows its neighbours need).  grid must be a multiple
// of 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t grid) {
    const uint32_t per = grid / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

// Own Cody-Waite exp for x <= 0 in fp32 (Cody and Waite, "Software Manual for
// the Elementary Functions", 1980): x = n*ln2 + r, |r| <= ln2/2, e^r by its
// degree-7 Taylor polynomial (coefficients 1/k!), times 2^n.  ln2 is split
// into a 16-bit head (0x3F317200) and the fp32-rounded remainder so n*head
// is exact.  Every operation is a single IEEE fp32 op or an fma, so host and
// device results are bit-identical.  Returns 0 below -86, where e^x < 2^-124
// (the caller compares a uniform in [0,1) against it); above, the result is
// a normal number, so denormal modes never matter.
VG_HD float exp_nonpos(float x) {
    if (x < -86.0f) return 0.0f;
    const float log2e = (float)1.4426950408889634;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = (float)(0.6931471805599453 - 0.693145751953125);
    const float n = rintf(x * log2e);
    float r = fmaf(n, -ln2_hi, x);
    r = fmaf(n, -ln2_lo, r);
    float p = (float)(1.0 / 5040.0);
    p = fmaf(p, r, (float)(1.0 / 720.0));
    p = fmaf(p, r, (float)(1.0 / 120.0));
    p = fmaf(p, r, (float)(1.0 / 24.0));
    p = fmaf(p, r, (float)(1.0 / 6.0));
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const int e = (int)n + 127;          // n in [-124, 0]: normal scale factor
    union {
        uint32_t u;
        float f;
    } s;
    s.u = (uint32_t)e << 23;
    return p * s.f;
}

// e^x for x <= 0 in f64, the graph chains' acceptance: Cody and Waite's
// reduction x = n ln2 + r (ln2 split so n * hi is exact, |r| <= ln2 / 2),
// then the degree-13 Taylor polynomial of e^r (coefficients 1/k! rounded to
// nearest), times 2^n; 0 below -708, where 2^n would leave the normal range.
// Every step is one IEEE operation or an fma, so host and device agree.
VG_HD double exp_f64_nonpos(double x) {
    if (x < -708.0) return 0.0;
    const double n = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(n, -0x1.62e42fee00000p-1, x);
    r = __builtin_fma(n, -0x1.a39ef35793c76p-33, r);
    double p = 0x1.6124613a86d09p-33;
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0x1.0000000000000p-1);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    union {
        uint64_t u;
        double d;
    } s;
    s.u = (uint64_t)((int64_t)n + 1023) << 52;
    return p * s.d;
}

// Correctly rounded fp32 square root for normal x > 0.  The hardware
// v_sqrt_f32 result s is within 1 ulp; with s- and s+ its neighbours, the
// exact residuals x - s-*s and x - s+*s (one fma each) decide whether the
// true root lies below the midpoint of s- and s, or above that of s and s+
// (a root of a float is never exactly a midpoint).  HIP's sqrtf and
// __fsqrt_rn are 1-ulp on gfx950, so they are not used where the CPU oracle
// must reproduce the bits.
__device__ __forceinline__ float sqrt_rn(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const uint32_t u = __float_as_uint(s);
    const float sm = __uint_as_float(u - 1u), sp = __uint_as_float(u + 1u);
    if (__builtin_fmaf(-sm, s, x) <= 0.0f)
        s = sm;
    else if (__builtin_fmaf(-sp, s, x) > 0.0f)
        s = sp;
    return s;
}

// Reductions over the 64 lanes of a wave (every lane active): DPP moves
// within each row of 16 lanes -- swap within pairs and within quads
// (quad_perm 1,0,3,2 and 2,3,0,1), then the half-row and row mirrors -- give
// every lane its row's total; the four rows' totals are read from lanes 0,
// 16, 32 and 48.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
    v += (int32_t)dpp<kDppXor1>((uint32_t)v);
    v += (int32_t)dpp<kDppXor2>((uint32_t)v);
    v += (int32_t)dpp<kDppHalfMirror>((uint32_t)v);
    v += (int32_t)dpp<kDppMirror>((uint32_t)v);
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) +
           __builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48);
}

// Issue priority of the calling wave, 0 (lowest) .. 3; p must fold to a
// constant.
__device__ __forceinline__ void set_prio(uint32_t p) {
    switch (p) {
        case 0: __builtin_amdgcn_s_setprio(0); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        default: __builtin_amdgcn_s_setprio(3); break;
    }
}

// Maximum over the 64 lanes of a wave (every lane active).
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    v = max(v, dpp<kDppXor1>(v));
    v = max(v, dpp<kDppXor2>(v));
    v = max(v, dpp<kDppHalfMirror>(v));
    v = max(v, dpp<kDppMirror>(v));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Ising colour planes are read and written through raw buffer resources:
// 32-bit offsets (a plane is below 4 GiB), and an offset past the plane
// (kNoRow) reads as zeros, which is exactly an absent neighbour row on an
// open boundary.  Dword 3 = 0x00020000 selects 32-bit data format on gfx9.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kNoRow = 0xFFFFFF00u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const uint8_t* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(p), 0, bytes, 0x00020000);
}

__device__ __forceinline__ u32x4 bload16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
}

__device__ __forceinline__ uint32_t bload8(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}

// 24-bit uniform in [0,1) and in [-1,1) from a 32-bit draw (exact in fp32).
VG_HD float unit24(uint32_t w) { return (float)(w >> 8) * 5.9604644775390625e-08f; }
VG_HD float sym24(uint32_t w) { return (float)(w >> 8) * 1.1920928955078125e-07f - 1.0f; }

}  // namespace vg
