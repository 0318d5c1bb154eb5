// Per-site fp32 Heisenberg update shared by the sc and bcc colour passes.
//
// The update restates MetropolisIntegrator::step (reference
// src/integrator.rs:66-92: fresh uniform-on-sphere proposal, Spin::rand ->
// marsaglia, util.rs:21-34) and MetropolisFlipIntegrator::step (109-138:
// proposal -s) with the local energy of Exchange + UniaxialAnisotropy +
// Zeeman + Gauge (src/energy.rs:64-214).  Every arithmetic step is a single
// IEEE fp32 operation or an explicit fused multiply-add, in a fixed order
// (no implicit contraction), so the CPU oracle reproduces every bit.
#pragma once

#include "vg_device.hpp"

namespace vg {

__device__ __forceinline__ float fm(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fa(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float fs(float a, float b) { return __fsub_rn(a, b); }
// Fused multiply-add, one rounding (v_fma_f32; the oracle's fmaf).
__device__ __forceinline__ float ff(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// a . b as one product and two fused multiply-adds, x then y then z.
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ff(az, bz, ff(ay, by, fm(ax, bx)));
}
__device__ __forceinline__ int64_t fix27(float v) {
    return (int64_t)__float2int_rn(fm(v, kFixScale));
}

__device__ __forceinline__ float4 ld4(const float* p, bool present) {
    if (!present) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return *reinterpret_cast<const float4*>(p);
}

// fp32 loads through a buffer resource (an offset past the plane reads 0).
__device__ __forceinline__ float4 bload4f(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                       __uint_as_float(v.w));
}
__device__ __forceinline__ float bload1f(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

__device__ __forceinline__ float f4get(const float4& v, uint32_t k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// 24-bit value m (0..4) of a Philox call: bits [24m, 24m+24) of the 128-bit
// string x | y << 32 | z << 64 | w << 96 (bits 120-127 unused).
__device__ __forceinline__ uint32_t v24(const U4& c, int m) {
    switch (m) {
        case 0: return c.x & 0xFFFFFFu;
        case 1: return __builtin_amdgcn_alignbit(c.y, c.x, 24) & 0xFFFFFFu;
        case 2: return __builtin_amdgcn_alignbit(c.z, c.y, 16) & 0xFFFFFFu;
        case 3: return __builtin_amdgcn_alignbit(c.w, c.z, 8) & 0xFFFFFFu;
        default: return c.w & 0xFFFFFFu;
    }
}
// A 24-bit value as a uniform in [0,1) and in [-1,1) (exact in fp32).
__device__ __forceinline__ float f24u(uint32_t v) { return fm((float)v, 5.9604644775390625e-08f); }
// (the product is exact, so the fused form rounds once, as fm then fs do)
__device__ __forceinline__ float f24s(uint32_t v) {
    return ff((float)v, 1.1920928955078125e-07f, -1.0f);
}

struct Try {
    float x1, x2, q;
};
__device__ __forceinline__ Try mtry(uint32_t a, uint32_t b) {
    Try t;
    t.x1 = f24s(a);
    t.x2 = f24s(b);
    t.q = fa(fm(t.x1, t.x1), fm(t.x2, t.x2));
    return t;
}

// A Marsaglia try inside the unit disc -> point on the sphere.
__device__ __forceinline__ void sphere_point(const Try& a, float p[3]) {
    const float root = sqrt_rn(fs(1.0f, a.q));
    p[0] = fm(fm(2.0f, a.x1), root);
    p[1] = fm(fm(2.0f, a.x2), root);
    p[2] = ff(-2.0f, a.q, 1.0f);  // 2q is exact: the same rounding as fs(1, 2q)
}

// Fresh uniform-sphere proposal of one site (Spin::rand -> marsaglia,
// util.rs:21-34) from its own Philox calls k = 0, 1, ... with counter
// (pos, row, sweep, purpose 16 + k): u = value 0 of call 0; Marsaglia tries
// (1,2), (3,4) of call 0, then (0,1), (2,3) of calls 1..15, the first inside
// the unit disc taken; exhausted -> (0,0,1).
//
// proposal_first does call 0 (both tries without branches) and returns
// whether one hit (p = 0.954); proposal_rest does calls 1..15 for a site
// whose first two tries missed.
__device__ __forceinline__ bool proposal_first(const ChainDev& C, uint32_t pos, uint32_t row,
                                               uint64_t t, uint32_t colour, float& u, float p[3]) {
    const U4 c0 = philox(U4{pos, row, (uint32_t)t, ctr_w(t, kHeisStream, colour)},
                         (uint32_t)C.seed, (uint32_t)(C.seed >> 32));
    u = f24u(v24(c0, 0));
    const Try t0 = mtry(v24(c0, 1), v24(c0, 2));
    const Try t1 = mtry(v24(c0, 3), v24(c0, 4));
    Try pick = t1;
    if (t0.q < 1.0f) pick = t0;
    const bool ok = pick.q < 1.0f;
    if (!ok) pick.q = 0.0f;  // keeps the square root finite; p is replaced later
    sphere_point(pick, p);
    return ok;
}

__device__ __forceinline__ void proposal_rest(const ChainDev& C, uint32_t pos, uint32_t row,
                                              uint64_t t, uint32_t colour, float p[3]) {
    const uint32_t k0 = (uint32_t)C.seed, k1 = (uint32_t)(C.seed >> 32);
    for (uint32_t k = 1; k < kHeisMaxCalls; ++k) {
        const U4 ck = philox(U4{pos, row, (uint32_t)t, ctr_w(t, kHeisStream + k, colour)}, k0, k1);
        const Try a = mtry(v24(ck, 0), v24(ck, 1));
        if (a.q < 1.0f) return sphere_point(a, p);
        const Try b = mtry(v24(ck, 2), v24(ck, 3));
        if (b.q < 1.0f) return sphere_point(b, p);
    }
    p[0] = 0.0f;
    p[1] = 0.0f;
    p[2] = 1.0f;
}

// Per-wave queue of the sites whose first call missed (about 3 in 64).
struct HeisQueue {
    uint32_t pos[64];
    uint32_t row[64];
    float p[3][64];
};

// Resolves the wave's pending sites: bit j of pend marks position i0 + j of
// this lane's row.  Entries are numbered by an exclusive prefix sum of the
// lanes' counts (three ballots); in rounds of 64, the owners post
// (position, row), lane e computes entry e's calls 1..15, and the owners
// read their proposals back.  LDS operations of one wave complete in order,
// so the wave needs no barrier, only no reordering by the compiler.  Every
// lane of the wave must call this.
__device__ __forceinline__ void heis_resolve(const ChainDev& C, HeisQueue& q, uint32_t pend,
                                             uint32_t i0, uint32_t row, uint64_t t,
                                             uint32_t colour, float p[4][3]) {
    const uint32_t lane = __lane_id();
    const uint32_t cnt = __builtin_popcount(pend);
    uint32_t pre = 0, total = 0;
#pragma unroll
    for (uint32_t bit = 0; bit < 3; ++bit) {
        const uint64_t m = __ballot((cnt >> bit) & 1u);
        pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))
               << bit;
        total += (uint32_t)__builtin_popcountll(m) << bit;
    }
    for (uint32_t base = 0; base < total; base += 64) {  // wave-uniform
        uint32_t slot = pre;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if ((pend >> j) & 1u) {
                const uint32_t e = slot - base;  // wraps past 64 when slot < base
                if (e < 64) {
                    q.pos[e] = i0 + j;
                    q.row[e] = row;
                }
                ++slot;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < total - base) {
            float r[3];
            proposal_rest(C, q.pos[lane], q.row[lane], t, colour, r);
            q.p[0][lane] = r[0];
            q.p[1][lane] = r[1];
            q.p[2][lane] = r[2];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        slot = pre;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if ((pend >> j) & 1u) {
                const uint32_t e = slot - base;
                if (e < 64) {
                    p[j][0] = q.p[0][e];
                    p[j][1] = q.p[1][e];
                    p[j][2] = q.p[2][e];
                }
                ++slot;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Flip proposals (-s) take their acceptance uniform from word (i & 3) of one
// call per 4 positions, counter (i >> 2, row, sweep, purpose 16).
__device__ __forceinline__ U4 flip_words(const ChainDev& C, uint32_t i0, uint32_t row, uint64_t t,
                                         uint32_t colour) {
    return philox(U4{i0 >> 2, row, (uint32_t)t, ctr_w(t, kHeisStream, colour)},
                  (uint32_t)C.seed, (uint32_t)(C.seed >> 32));
}

__device__ __forceinline__ int32_t fix27_32(float v) { return __float2int_rn(fm(v, kFixScale)); }

// One site: old spin o, neighbour sum h, acceptance uniform u, proposal p.
// Writes the new spin to n and adds the site's fixed-point sums: mx, my, mz
// and (s.axis)^2 to s32 (each at most 2^27 in magnitude per site, so a lane
// may collect up to 15 sites in int32) and, when `bond` is set, n.hb to
// bsum (hb: the sum over the neighbours of lower colours; for the
// two-colour lattices every neighbour of a colour-1 site).  A site that is
// not `valid` (zero padding) keeps o and adds nothing.
__device__ __forceinline__ void heis_update(const ChainDev& C, bool bond, const float o[3],
                                            const float h[3], const float hb[3], float u,
                                            const float p[3], float n[3], int32_t s32[4],
                                            int64_t& bsum, bool valid = true) {
    const float ox = o[0], oy = o[1], oz = o[2];
    const float px = p[0], py = p[1], pz = p[2];
    const float ax = C.ax, ay = C.ay, az = C.az;
    const float dx = fs(px, ox), dy = fs(py, oy), dz = fs(pz, oz);
    const float e_ex = fm(C.jn, dot3(dx, dy, dz, h[0], h[1], h[2]));
    const float pe = dot3(px, py, pz, ax, ay, az);
    const float oe = dot3(ox, oy, oz, ax, ay, az);
    const float e_an = fm(C.kk, ff(pe, pe, -fm(oe, oe)));
    const float e_ze = fm(C.hm, dot3(dx, dy, dz, C.fx, C.fy, C.fz));
    const float dE = fa(fa(e_ex, e_an), e_ze);
    bool accept = valid && dE < 0.0f;
    if (valid && !accept) accept = u < exp_nonpos(-fm(dE, C.beta));
    n[0] = accept ? px : ox;
    n[1] = accept ? py : oy;
    n[2] = accept ? pz : oz;
    s32[0] += fix27_32(n[0]);
    s32[1] += fix27_32(n[1]);
    s32[2] += fix27_32(n[2]);
    const float se = accept ? pe : oe;  // n.axis, the same operations as dot3(n, a)
    s32[3] += fix27_32(fm(se, se));
    if (bond) bsum += fix27(dot3(n[0], n[1], n[2], hb[0], hb[1], hb[2]));
}

// The int32 site sums of a lane into the five int64 quantities.
__device__ __forceinline__ void heis_fold(const int32_t s32[4], int64_t bsum, int64_t acc[5]) {
    acc[0] += bsum;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[1 + k] += (int64_t)s32[k];
}

// Phase 1 of a site: its acceptance uniform u and proposal p.  Random
// proposals whose first call missed set bit j of pend (resolved by
// heis_resolve); flip proposals take word j of fw.
__device__ __forceinline__ void heis_propose(const ChainDev& C, uint32_t colour, uint32_t pos,
                                             uint32_t row, uint64_t t, const U4& fw, uint32_t j,
                                             const float o[3], float& u, float p[3],
                                             uint32_t& pend) {
    if (C.proposal) {
        if (!proposal_first(C, pos, row, t, colour, u, p)) pend |= 1u << j;
    } else {
        u = unit24(u4_word(fw, j));
        p[0] = -o[0];
        p[1] = -o[1];
        p[2] = -o[2];
    }
}

// The workgroup's sums of the five quantities, from each wave's totals w
// (equal in every lane) -> partials[sweep][colour][block][5] (ncol colours
// per sweep).
__device__ __forceinline__ void heis_block_partials(const int64_t w[5],
                                                    int64_t* __restrict__ partials,
                                                    uint32_t colour, uint32_t ncol, uint32_t s,
                                                    uint32_t total_blocks, uint32_t b) {
    __shared__ int64_t red[5][kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) red[k][wave] = w[k];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        int64_t v = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBlock / 64; ++k) v += red[threadIdx.x][k];
        partials[((uint64_t)(s * ncol + colour) * total_blocks + b) * 5u + threadIdx.x] = v;
    }
}

// Workgroup sums of the five quantities of any magnitude (the graph passes,
// whose couplings enter the bond share): 64-bit shuffle reduction.
__device__ __forceinline__ void heis_partials(int64_t acc[5], int64_t* __restrict__ partials,
                                              uint32_t colour, uint32_t ncol, uint32_t s,
                                              uint32_t total_blocks, uint32_t b) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    }
    heis_block_partials(acc, partials, colour, ncol, s, total_blocks, b);
}

// Sum over the wave (every lane active) of a value below 2^40 in magnitude
// in every lane: its low 16 bits and the rest (below 2^24) are summed in
// int32 -- below 2^22 and 2^30 over 64 lanes, so exactly.
__device__ __forceinline__ int64_t wave_sum_split(int64_t v) {
    const int32_t lo = (int32_t)(v & 0xFFFF);
    const int32_t hi = (int32_t)(v >> 16);
    return (int64_t)wave_sum(hi) * 65536 + (int64_t)wave_sum(lo);
}

// Workgroup sums of the lattice passes (sc, bcc, fcc): a lane holds at most
// four sites, each below 2^27 in every magnetisation sum and below
// z * 2^27 <= 12 * 2^27 in its bond share, so every lane's five values are
// below 2^40 and reduce with wave_sum_split.
__device__ __forceinline__ void heis_partials_sites(const int32_t s32[4], int64_t bsum,
                                                    int64_t* __restrict__ partials,
                                                    uint32_t colour, uint32_t ncol, uint32_t s,
                                                    uint32_t total_blocks, uint32_t b) {
    int64_t w[5];
    w[0] = wave_sum_split(bsum);
#pragma unroll
    for (int k = 0; k < 4; ++k) w[1 + k] = wave_sum_split(s32[k]);
    heis_block_partials(w, partials, colour, ncol, s, total_blocks, b);
}

}  // namespace vg
