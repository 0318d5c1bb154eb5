// Per-site fp32 Heisenberg update shared by the sc and bcc colour passes.
//
// The update restates MetropolisIntegrator::step (reference
// src/integrator.rs:66-92: fresh uniform-on-sphere proposal, Spin::rand ->
// marsaglia, util.rs:21-34) and MetropolisFlipIntegrator::step (109-138:
// proposal -s) with the local energy of Exchange + UniaxialAnisotropy +
// Zeeman + Gauge (src/energy.rs:64-214).  Every arithmetic step is a single
// IEEE fp32 operation in a fixed order (no contraction), so the CPU oracle
// reproduces every bit.
#pragma once

#include "vg_device.hpp"

namespace vg {

__device__ __forceinline__ float fm(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fa(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float fs(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return fa(fa(fm(ax, bx), fm(ay, by)), fm(az, bz));
}
__device__ __forceinline__ int64_t fix27(float v) {
    return (int64_t)__float2int_rn(fm(v, kFixScale));
}

__device__ __forceinline__ float4 ld4(const float* p, bool present) {
    if (!present) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return *reinterpret_cast<const float4*>(p);
}

__device__ __forceinline__ float f4get(const float4& v, uint32_t k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// Sequential word stream of one 4-position group: Philox counter
// (i/4, row, sweep, purpose 16 + call), consumed in position order.
struct Stream {
    U4 cur;
    uint32_t call, word;
    uint32_t cx, cy, cz, sweep_hi_bits, k0, k1;
    __device__ __forceinline__ void start(const ChainDev& C, uint32_t i0, uint32_t row,
                                          uint64_t t, uint32_t colour) {
        call = 0;
        cx = i0 >> 2;
        cy = row;
        cz = (uint32_t)t;
        sweep_hi_bits = ((uint32_t)(t >> 32) << 8) | colour;
        k0 = (uint32_t)C.seed;
        k1 = (uint32_t)(C.seed >> 32);
        refill();
    }
    __device__ __forceinline__ void refill() {
        cur = philox(U4{cx, cy, cz, sweep_hi_bits | ((kHeisStream + call) << 1)}, k0, k1);
        word = 0;
    }
    // Returns false once the stream is exhausted (after kHeisMaxCalls calls).
    __device__ __forceinline__ bool next(uint32_t& out) {
        if (word == 4) {
            if (call + 1 >= kHeisMaxCalls) return false;
            ++call;
            refill();
        }
        out = u4_word(cur, word++);
        return true;
    }
};

// One site: old spin o, neighbour sum h.  Draws its acceptance uniform (one
// word), then for a random proposal two words per Marsaglia try until one
// lands inside the unit disc; an exhausted stream gives u = 0 and proposal
// (0,0,1).  Returns the new spin in n and adds the site's fixed-point sums
// {bond (colour 1 only), mx, my, mz, (s.axis)^2} to acc.
__device__ __forceinline__ void heis_site(const ChainDev& C, Stream& st, bool& live,
                                          uint32_t colour, const float o[3], const float h[3],
                                          float n[3], int64_t acc[5]) {
    const float ox = o[0], oy = o[1], oz = o[2];
    uint32_t wu = 0;
    float px = -ox, py = -oy, pz = -oz;
    live = live && st.next(wu);
    const float u = unit24(wu);
    if (C.proposal) {
        px = 0.0f;
        py = 0.0f;
        pz = 1.0f;
        while (live) {
            uint32_t w1, w2;
            if (!st.next(w1) || !st.next(w2)) {
                live = false;
                break;
            }
            const float x1 = sym24(w1), x2 = sym24(w2);
            const float qq = fa(fm(x1, x1), fm(x2, x2));
            if (qq < 1.0f) {
                const float root = sqrt_rn(fs(1.0f, qq));
                px = fm(fm(2.0f, x1), root);
                py = fm(fm(2.0f, x2), root);
                pz = fs(1.0f, fm(2.0f, qq));
                break;
            }
        }
    }
    const float ax = C.ax, ay = C.ay, az = C.az;
    const float dx = fs(px, ox), dy = fs(py, oy), dz = fs(pz, oz);
    const float e_ex = fm(C.jn, dot3(dx, dy, dz, h[0], h[1], h[2]));
    const float pe = dot3(px, py, pz, ax, ay, az);
    const float oe = dot3(ox, oy, oz, ax, ay, az);
    const float e_an = fm(C.kk, fs(fm(pe, pe), fm(oe, oe)));
    const float e_ze = fm(C.hm, dot3(dx, dy, dz, C.fx, C.fy, C.fz));
    const float dE = fa(fa(e_ex, e_an), e_ze);
    bool accept = dE < 0.0f;
    if (!accept) accept = u < exp_nonpos(-fm(dE, C.beta));
    n[0] = accept ? px : ox;
    n[1] = accept ? py : oy;
    n[2] = accept ? pz : oz;
    acc[1] += fix27(n[0]);
    acc[2] += fix27(n[1]);
    acc[3] += fix27(n[2]);
    const float se = dot3(n[0], n[1], n[2], ax, ay, az);
    acc[4] += fix27(fm(se, se));
    if (colour == 1) acc[0] += fix27(dot3(n[0], n[1], n[2], h[0], h[1], h[2]));
}

// Workgroup sums of the five quantities -> partials[sweep][colour][block][5].
__device__ __forceinline__ void heis_partials(int64_t acc[5], int64_t* __restrict__ partials,
                                              uint32_t colour, uint32_t s, uint32_t total_blocks,
                                              uint32_t b) {
    __shared__ int64_t red[5][kBlock / 64];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) red[k][wave] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        int64_t v = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBlock / 64; ++k) v += red[threadIdx.x][k];
        partials[((uint64_t)(s * 2u + colour) * total_blocks + b) * 5u + threadIdx.x] = v;
    }
}

}  // namespace vg
