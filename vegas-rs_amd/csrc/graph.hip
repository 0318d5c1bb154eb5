// Colour-class sweeps of chains whose exchange is an arbitrary sparse matrix
// J: Exchange::new(CsMat<f64>) of the reference (src/energy.rs:170-172),
// with e_i = -sum_j J_ij s_i.s_j over row i (energy.rs:194-206) inside the
// Metropolis step of MetropolisFlipIntegrator / MetropolisIntegrator
// (src/integrator.rs:66-138).
//
// The host colours the symmetrised pattern greedily and stores the sites of
// a chain sorted by (colour, id); a colour class has no two sites that share
// a nonzero, so one pass updates a class all at once, one thread per site.
// Each position's merged row lists its neighbours (ascending site id) with
// a = J_ij and w = J_ij + J_ji for neighbours of lower colour (0 otherwise):
// h = sum a_j s_j is the field of the move, and s_i.(sum w_j s_j) after the
// move is the site's share of sum_{i != j} J_ij s_i.s_j, every pair counted
// at its higher-coloured end when both spins are final for the sweep.
//
// Ising: f64, every operation rounded on its own (no contraction), the
// acceptance `dE <= 0 or u < exp(-dE/T)` with u the 32-bit draw times
// 2^-32 and the engine's own exp (exp_f64_nonpos).  Heisenberg: the fp32
// update of the lattice passes with the couplings inside h.  Draws are keyed
// by the site id, so the CPU restatement reproduces every bit.
//
// Memory: each site reads its row (4 + 2 x 8 bytes per entry for Ising, 4 +
// 2 x 4 for Heisenberg) and its neighbours' spins by gather; this path is
// for generality (any lattice or coupling set), the stencil passes are the
// fast path for sc, bcc and fcc.
#include "heis_common.hpp"

namespace vg {

__global__ __launch_bounds__(256) void graph_ising_pass(const ChainDev* __restrict__ chains,
                                                        const uint32_t* __restrict__ block_chain,
                                                        uint8_t* __restrict__ arena,
                                                        int64_t* __restrict__ partials,
                                                        uint32_t colour, uint32_t s,
                                                        uint32_t total_blocks, uint32_t npass) {
    const uint32_t b = blockIdx.x;
    const ChainDev& C = chains[block_chain[b]];
    if (s >= C.limit) return;  // uniform: the workgroup belongs to one chain
    int64_t acc[5] = {0, 0, 0, 0, 0};
    if (colour < C.ncol) {
        const uint32_t* cst = reinterpret_cast<const uint32_t*>(arena + C.goff[2]);
        const uint32_t p = cst[colour] + (b - C.block0) * kBlock + threadIdx.x;
        if (p < cst[colour + 1]) {
            uint8_t* spin = arena + C.off[0];
            const uint32_t* ptr = reinterpret_cast<const uint32_t*>(arena + C.off[1]);
            const uint32_t* nbr = reinterpret_cast<const uint32_t*>(arena + C.off[2]);
            const double* a = reinterpret_cast<const double*>(arena + C.off[3]);
            const double* w = reinterpret_cast<const double*>(arena + C.goff[0]);
            const uint32_t site = reinterpret_cast<const uint32_t*>(arena + C.goff[1])[p];
            const uint64_t t = C.sweep0 + s;
            const U4 r = philox(U4{site, 0u, (uint32_t)t, ctr_w(t, kGraphIsing, 0)},
                                (uint32_t)C.seed, (uint32_t)(C.seed >> 32));
            double sv = spin[p] ? 1.0 : -1.0;
            // Spin::rand proposes Up when its draw is below 1/2 (state.rs:76-84);
            // a proposal equal to the current spin changes nothing.
            const bool change = C.proposal == 0 || ((r.y < 0x80000000u) ? 1.0 : -1.0) != sv;
            double h = 0.0, hb = 0.0;
            const uint32_t e1 = ptr[p + 1];
            for (uint32_t e = ptr[p]; e < e1; ++e) {
                const double sj = spin[nbr[e]] ? 1.0 : -1.0;
                h = __dadd_rn(h, __dmul_rn(a[e], sj));
                hb = __dadd_rn(hb, __dmul_rn(w[e], sj));
            }
            // The decision is computed without branches and applied by one
            // select: written as `if (dE <= 0 || u < p) flip`, the accept
            // block has two predecessors, and hipcc for gfx950 (ROCm 7.2)
            // kept the old spin on the second path (the optimised IR is
            // right).  exp is evaluated for every site; its argument is
            // clamped to <= 0, which changes nothing where it is used
            // (dE > 0).
            const double dE = __dsub_rn(__dmul_rn(__dmul_rn(2.0, sv), h),
                                        __dmul_rn(__dmul_rn(__dmul_rn(2.0, C.ghm), C.ghz), sv));
            const double pr = exp_f64_nonpos(fmin(-__dmul_rn(dE, C.gbeta), 0.0));
            const bool accept =
                change & ((dE <= 0.0) | (__dmul_rn((double)r.x, 0x1p-32) < pr));
            if (accept) spin[p] = sv > 0.0 ? 0 : 1;
            sv = accept ? -sv : sv;
            acc[3] = sv > 0.0 ? 1 : -1;
            acc[0] = __double2ll_rn(__dmul_rn(__dmul_rn(sv, hb), 134217728.0));
        }
    }
    heis_partials(acc, partials, colour, npass, s, total_blocks, b);
}

__global__ __launch_bounds__(256) void graph_heis_pass(const ChainDev* __restrict__ chains,
                                                       const uint32_t* __restrict__ block_chain,
                                                       uint8_t* __restrict__ arena,
                                                       int64_t* __restrict__ partials,
                                                       uint32_t colour, uint32_t s,
                                                       uint32_t total_blocks, uint32_t npass) {
    const uint32_t b = blockIdx.x;
    const ChainDev& C = chains[block_chain[b]];
    if (s >= C.limit) return;
    int64_t acc[5] = {0, 0, 0, 0, 0};
    if (colour < C.ncol) {
        const uint32_t* cst = reinterpret_cast<const uint32_t*>(arena + C.goff[2]);
        const uint32_t p = cst[colour] + (b - C.block0) * kBlock + threadIdx.x;
        if (p < cst[colour + 1]) {
            float* xs = reinterpret_cast<float*>(arena + C.off[0]);
            float* ys = xs + C.nsites;
            float* zs = ys + C.nsites;
            const uint32_t* ptr = reinterpret_cast<const uint32_t*>(arena + C.off[1]);
            const uint32_t* nbr = reinterpret_cast<const uint32_t*>(arena + C.off[2]);
            const float* a = reinterpret_cast<const float*>(arena + C.off[3]);
            const float* w = reinterpret_cast<const float*>(arena + C.goff[0]);
            const uint32_t site = reinterpret_cast<const uint32_t*>(arena + C.goff[1])[p];
            const uint64_t t = C.sweep0 + s;
            float h[3] = {0.0f, 0.0f, 0.0f}, hb[3] = {0.0f, 0.0f, 0.0f};
            const uint32_t e1 = ptr[p + 1];
            for (uint32_t e = ptr[p]; e < e1; ++e) {
                const uint32_t j = nbr[e];
                const float aj = a[e], wj = w[e];
                const float sx = xs[j], sy = ys[j], sz = zs[j];
                h[0] = fa(h[0], fm(aj, sx));
                h[1] = fa(h[1], fm(aj, sy));
                h[2] = fa(h[2], fm(aj, sz));
                hb[0] = fa(hb[0], fm(wj, sx));
                hb[1] = fa(hb[1], fm(wj, sy));
                hb[2] = fa(hb[2], fm(wj, sz));
            }
            const float o[3] = {xs[p], ys[p], zs[p]};
            float u, pr[3];
            if (C.proposal) {
                if (!proposal_first(C, site, 0u, t, 0u, u, pr)) proposal_rest(C, site, 0u, t, 0u, pr);
            } else {
                const U4 fw = philox(U4{site, 0u, (uint32_t)t, ctr_w(t, kHeisStream, 0)},
                                     (uint32_t)C.seed, (uint32_t)(C.seed >> 32));
                u = unit24(fw.x);
                pr[0] = -o[0];
                pr[1] = -o[1];
                pr[2] = -o[2];
            }
            float n[3];
            int32_t s32[4] = {0, 0, 0, 0};
            int64_t bsum = 0;
            heis_update(C, false, o, h, hb, u, pr, n, s32, bsum);
            heis_fold(s32, bsum, acc);
            // the bond share can exceed the int32 fixed-point range of the
            // lattice sums when couplings are large: 64-bit conversion
            acc[0] = (int64_t)__float2ll_rn(fm(dot3(n[0], n[1], n[2], hb[0], hb[1], hb[2]), kFixScale));
            xs[p] = n[0];
            ys[p] = n[1];
            zs[p] = n[2];
        }
    }
    heis_partials(acc, partials, colour, npass, s, total_blocks, b);
}

// Initial states (State::rand_with_size, src/state.rs:260-262) keyed by the
// site id: Ising Up iff word 0 of Philox(site, 0, 0, w(0, graph-init, 0)) is
// below 2^31; Heisenberg by Marsaglia's method on the calls
// Philox(site, 0, k, w(0, heis-init, 0)), two tries per call, (0,0,1) after
// 64 calls.  One thread per position.
__global__ void graph_ising_init(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C.nsites) return;
    const uint32_t site = reinterpret_cast<const uint32_t*>(arena + C.goff[1])[p];
    const U4 r = philox(U4{site, 0u, 0u, ctr_w(0, kGraphInit, 0)}, (uint32_t)C.seed,
                        (uint32_t)(C.seed >> 32));
    arena[C.off[0] + p] = r.x < 0x80000000u ? 1 : 0;
}

__global__ void graph_heis_init(const ChainDev* __restrict__ chain, uint8_t* __restrict__ arena) {
    const ChainDev& C = *chain;
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= C.nsites) return;
    const uint32_t site = reinterpret_cast<const uint32_t*>(arena + C.goff[1])[p];
    float sx = 0.0f, sy = 0.0f, sz = 1.0f;
    for (uint32_t k = 0; k < kHeisInitMaxCalls; ++k) {
        const U4 r = philox(U4{site, 0u, k, ctr_w(0, kHeisInit, 0)}, (uint32_t)C.seed,
                            (uint32_t)(C.seed >> 32));
        bool done = false;
#pragma unroll
        for (uint32_t q = 0; q < 2; ++q) {
            const float x1 = sym24(u4_word(r, 2 * q));
            const float x2 = sym24(u4_word(r, 2 * q + 1));
            const float qq = fa(fm(x1, x1), fm(x2, x2));
            if (!done && qq < 1.0f) {
                const float root = sqrt_rn(fs(1.0f, qq));
                sx = fm(fm(2.0f, x1), root);
                sy = fm(fm(2.0f, x2), root);
                sz = fs(1.0f, fm(2.0f, qq));
                done = true;
            }
        }
        if (done) break;
    }
    float* xs = reinterpret_cast<float*>(arena + C.off[0]);
    xs[p] = sx;
    xs[C.nsites + p] = sy;
    xs[2 * C.nsites + p] = sz;
}

}  // namespace vg
