// Layout kernels: the i.i.d. initial state (State::rand_with_size,
// reference src/state.rs:76-84, 145-148, 260-262) drawn from each chain's
// Philox stream, and conversion between the natural site order used at the
// C ABI and the colour-split planes the sweep kernels use.
#include "vg_device.hpp"

namespace vg {

// Natural site id -> (colour, row, position): sc site x + lx*(y + ly*z);
// bcc and fcc site (x + lx*(y + ly*z)) * k + sublattice, k = 2 or 4.
struct Where {
    uint32_t c, row, i;
};

__device__ __forceinline__ Where locate(const ChainDev& C, uint64_t n) {
    Where r;
    if (C.kind == 0) {
        const uint32_t x = (uint32_t)(n % C.lx);
        const uint64_t yz = n / C.lx;
        const uint32_t y = (uint32_t)(yz % C.ly);
        const uint32_t z = (uint32_t)(yz / C.ly);
        r.c = (x + y + z) & 1u;
        r.row = y + C.ly * z;
        r.i = x >> 1;
    } else {
        const uint32_t sh = C.ncol == 4 ? 2u : 1u;
        const uint64_t cell = n >> sh;
        const uint32_t x = (uint32_t)(cell % C.lx);
        const uint64_t yz = cell / C.lx;
        const uint32_t y = (uint32_t)(yz % C.ly);
        const uint32_t z = (uint32_t)(yz / C.ly);
        r.c = (uint32_t)(n & ((1u << sh) - 1u));
        r.row = y + C.ly * z;
        r.i = x;
    }
    return r;
}

// Ising initial state: position i of row `row` in plane c is Up iff word
// (i & 3) of Philox(key = seed, ctr = (i >> 2, row, 0, w(0, init, c))) is
// below 2^31 (Spin::rand draws Up when its uniform is below 1/2).
__global__ void ising_init(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per_plane = (uint64_t)C.rows * C.pitch;
    if (idx >= C.ncol * per_plane) return;
    const uint32_t c = (uint32_t)(idx / per_plane);
    const uint64_t e = idx - c * per_plane;
    const uint32_t row = (uint32_t)(e / C.pitch);
    const uint32_t i = (uint32_t)(e % C.pitch);
    uint8_t v = 0;
    if (i < C.w) {
        const U4 r = philox(U4{i >> 2, row, 0u, ctr_w(0, kIsingInit, c)}, (uint32_t)C.seed,
                            (uint32_t)(C.seed >> 32));
        v = (u4_word(r, i & 3u) >> 31) == 0 ? 1 : 0;
    }
    arena[C.off[c] + e] = v;
}

__global__ void ising_scatter(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena,
                              const int8_t* __restrict__ natural, uint64_t n_sites) {
    const ChainDev& C = *chain;
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_sites) return;
    const Where p = locate(C, n);
    arena[C.off[p.c] + (uint64_t)p.row * C.pitch + p.i] = natural[n] > 0 ? 1 : 0;
}

__global__ void ising_gather(const ChainDev* __restrict__ chain, const uint8_t* __restrict__ arena,
                             int8_t* __restrict__ natural, uint64_t n_sites) {
    const ChainDev& C = *chain;
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_sites) return;
    const Where p = locate(C, n);
    natural[n] = arena[C.off[p.c] + (uint64_t)p.row * C.pitch + p.i] ? 1 : -1;
}

// Bit-packed Ising planes: position i of a row is bit (i & 7) of byte i >> 3.
// One thread per plane byte; bits past w stay zero.

// Natural site id of position i of row `row` of colour c: sc x + lx*(y +
// ly*z) with x = 2i + ((c+y+z)&1); bcc and fcc ncol*(cell) + c with the
// cell's x = i.
__device__ __forceinline__ uint64_t plane_site(const ChainDev& C, uint32_t c, uint32_t row,
                                               uint32_t i) {
    const uint32_t y = row % C.ly, z = row / C.ly;
    if (C.kind == 0) {
        const uint32_t x = 2u * i + ((c + y + z) & 1u);
        return (uint64_t)x + (uint64_t)C.lx * ((uint64_t)y + (uint64_t)C.ly * z);
    }
    return (uint64_t)C.ncol * ((uint64_t)i + (uint64_t)C.lx * ((uint64_t)y + (uint64_t)C.ly * z)) +
           c;
}

// The same draws as ising_init: position i is Up iff word (i & 3) of
// Philox(key = seed, ctr = (i >> 2, row, 0, w(0, init, c))) is below 2^31.
__global__ void ising_init_bits(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per_plane = (uint64_t)C.rows * C.pitch;
    if (idx >= C.ncol * per_plane) return;
    const uint32_t c = (uint32_t)(idx / per_plane);
    const uint64_t e = idx - c * per_plane;
    const uint32_t row = (uint32_t)(e / C.pitch);
    const uint32_t i0 = 8u * (uint32_t)(e % C.pitch);
    uint32_t v = 0;
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t i = i0 + 4u * h;
        if (i >= C.w) break;
        const U4 r = philox(U4{i >> 2, row, 0u, ctr_w(0, kIsingInit, c)}, (uint32_t)C.seed,
                            (uint32_t)(C.seed >> 32));
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            if (i + k < C.w && (u4_word(r, k) >> 31) == 0) v |= 1u << (4u * h + k);
    }
    arena[C.off[c] + e] = (uint8_t)v;
}

__global__ void ising_scatter_bits(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena,
                                   const int8_t* __restrict__ natural) {
    const ChainDev& C = *chain;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per_plane = (uint64_t)C.rows * C.pitch;
    if (idx >= C.ncol * per_plane) return;
    const uint32_t c = (uint32_t)(idx / per_plane);
    const uint64_t e = idx - c * per_plane;
    const uint32_t row = (uint32_t)(e / C.pitch);
    const uint32_t i0 = 8u * (uint32_t)(e % C.pitch);
    uint32_t v = 0;
    for (uint32_t k = 0; k < 8 && i0 + k < C.w; ++k)
        if (natural[plane_site(C, c, row, i0 + k)] > 0) v |= 1u << k;
    arena[C.off[c] + e] = (uint8_t)v;
}

__global__ void ising_gather_bits(const ChainDev* __restrict__ chain,
                                  const uint8_t* __restrict__ arena, int8_t* __restrict__ natural,
                                  uint64_t n_sites) {
    const ChainDev& C = *chain;
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_sites) return;
    const Where p = locate(C, n);
    const uint8_t v = arena[C.off[p.c] + (uint64_t)p.row * C.pitch + (p.i >> 3)];
    natural[n] = (v >> (p.i & 7u)) & 1u ? 1 : -1;
}

// Zero the padding of a Heisenberg chain (positions >= w) and, before any
// state is set, everything.
__global__ void heis_clear(const ChainDev* __restrict__ chain, float* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 3ull * C.ncol * C.plane) return;
    const uint32_t c = (uint32_t)(idx / (3 * C.plane));
    const uint64_t e = idx - (uint64_t)c * 3 * C.plane;
    arena[C.off[c] + e] = 0.0f;
}

// Heisenberg initial state: Marsaglia (1972) on S^2 (reference util.rs:21-34)
// in fp32.  Position i of row `row` in plane c consumes the words of
// Philox(key = seed, ctr = (i, row, k, w(0, heis-init, c))), k = 0, 1, ...,
// two per try: x1 = sym24(word 2t), x2 = sym24(word 2t+1); reject while
// x1^2 + x2^2 >= 1.  After 64 calls without success the spin is (0,0,1).
__global__ void heis_init(const ChainDev* __restrict__ chain, float* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per_colour = (uint64_t)C.rows * C.w;
    if (idx >= C.ncol * per_colour) return;
    const uint32_t c = (uint32_t)(idx / per_colour);
    const uint64_t e = idx - c * per_colour;
    const uint32_t row = (uint32_t)(e / C.w);
    const uint32_t i = (uint32_t)(e % C.w);
    float sx = 0.0f, sy = 0.0f, sz = 1.0f;
    for (uint32_t k = 0; k < kHeisInitMaxCalls; ++k) {
        const U4 r = philox(U4{i, row, k, ctr_w(0, kHeisInit, c)}, (uint32_t)C.seed,
                            (uint32_t)(C.seed >> 32));
        bool done = false;
#pragma unroll
        for (uint32_t t = 0; t < 2; ++t) {
            const float x1 = sym24(u4_word(r, 2 * t));
            const float x2 = sym24(u4_word(r, 2 * t + 1));
            const float q = __fadd_rn(__fmul_rn(x1, x1), __fmul_rn(x2, x2));
            if (!done && q < 1.0f) {
                const float root = sqrt_rn(__fsub_rn(1.0f, q));
                sx = __fmul_rn(__fmul_rn(2.0f, x1), root);
                sy = __fmul_rn(__fmul_rn(2.0f, x2), root);
                sz = __fsub_rn(1.0f, __fmul_rn(2.0f, q));
                done = true;
            }
        }
        if (done) break;
    }
    const uint64_t o = C.off[c] + (uint64_t)row * C.pitch + i;
    arena[o] = sx;
    arena[o + C.plane] = sy;
    arena[o + 2 * C.plane] = sz;
}

__global__ void heis_scatter(const ChainDev* __restrict__ chain, float* __restrict__ arena,
                             const double* __restrict__ xyz, uint64_t n_sites) {
    const ChainDev& C = *chain;
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_sites) return;
    const Where p = locate(C, n);
    const uint64_t o = C.off[p.c] + (uint64_t)p.row * C.pitch + p.i;
    arena[o] = __double2float_rn(xyz[3 * n + 0]);
    arena[o + C.plane] = __double2float_rn(xyz[3 * n + 1]);
    arena[o + 2 * C.plane] = __double2float_rn(xyz[3 * n + 2]);
}

__global__ void heis_gather(const ChainDev* __restrict__ chain, const float* __restrict__ arena,
                            double* __restrict__ xyz, uint64_t n_sites) {
    const ChainDev& C = *chain;
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_sites) return;
    const Where p = locate(C, n);
    const uint64_t o = C.off[p.c] + (uint64_t)p.row * C.pitch + p.i;
    xyz[3 * n + 0] = (double)arena[o];
    xyz[3 * n + 1] = (double)arena[o + C.plane];
    xyz[3 * n + 2] = (double)arena[o + 2 * C.plane];
}

// Zero n int64 words (used for per-block records; a kernel rather than a
// memset so the sweep block can be captured in a graph).
__global__ void zero_i64(int64_t* __restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace vg
