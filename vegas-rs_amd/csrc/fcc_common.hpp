// Face-centred cubic geometry shared by the fcc colour passes.
//
// The four sublattices of the fcc cell sit at (0,0,0), (0,1/2,1/2),
// (1/2,0,1/2) and (1/2,1/2,0); no two sites of one sublattice are nearest
// neighbours, so the sublattices are the colours of a four-colour sweep.  A
// site of sublattice a in cell c bonds to four sites of each other sublattice
// b: those in cells c + d1*e1 + d2*e2, d1, d2 in {0,1}, over the two axes e1 <
// e2 along which a and b differ, with e = -1 along an axis where b sits
// higher in its cell than a and +1 where lower.  Slot order (the order the
// Heisenberg field is summed in): b ascending, then (d1, d2) = (0,0), (1,0),
// (0,1), (1,1).
//
// Along x, sublattices 0 and 1 sit at 0, sublattices 2 and 3 at 1/2: every
// x-offset neighbour of a colour-0 or colour-1 site is at x-1, of a colour-2
// or colour-3 site at x+1.
#pragma once

#include "vg_device.hpp"

namespace vg {

// Bit e (x, y, z) of sublattice s: its half-cell offset along axis e.
__host__ __device__ constexpr uint32_t fcc_half(uint32_t s, uint32_t e) {
    return ((s == 0 ? 0u : (s == 1 ? 6u : (s == 2 ? 5u : 3u))) >> e) & 1u;
}
// Axes (bit mask) along which sublattices a and b differ: two of x, y, z.
__host__ __device__ constexpr uint32_t fcc_diff(uint32_t a, uint32_t b) {
    return (fcc_half(a, 0) ^ fcc_half(b, 0)) | ((fcc_half(a, 1) ^ fcc_half(b, 1)) << 1) |
           ((fcc_half(a, 2) ^ fcc_half(b, 2)) << 2);
}
// Step (+1 or -1) towards b's neighbours of an a site along axis e.
__host__ __device__ constexpr int fcc_step(uint32_t a, uint32_t b, uint32_t e) {
    return fcc_half(b, e) > fcc_half(a, e) ? -1 : 1;
}

// Coordinate v + d (d = +1 or -1) on an axis of length n: wraps when
// periodic; ok = false when it leaves an open axis.
__device__ __forceinline__ uint32_t fcc_move(uint32_t v, int d, uint32_t n, bool periodic,
                                             bool& ok) {
    if (d < 0) {
        ok = v > 0 || periodic;
        return v > 0 ? v - 1 : n - 1;
    }
    ok = v + 1 < n || periodic;
    return v + 1 < n ? v + 1 : 0;
}

}  // namespace vg
