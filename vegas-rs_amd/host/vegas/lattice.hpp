// Lattice topology for the Hamiltonian's Exchange term: the unit cells and
// operations the reference takes from the vegas-lattice crate
// (src/input.rs:296-322, src/main.rs:25): Lattice::sc / bcc / fcc(a),
// expand(x,y,z), expand_all(n), expand_x/y/z(n), drop_x/y/z() (periodic ->
// open).  Site ids: sc x + lx*(y + ly*z); bcc 2*(cell) + sublattice; fcc
// 4*(cell) + sublattice (corner, then the face centres (0,1/2,1/2),
// (1/2,0,1/2), (1/2,1/2,0)).  The engine sweeps sc and bcc in two colours and
// fcc in four, one per sublattice.
#pragma once

#include <cstdint>
#include <vector>

#include "error.hpp"

namespace vegas {

class Lattice {
  public:
    enum Cell { SC = VG_LATTICE_SC, BCC = VG_LATTICE_BCC, FCC = VG_LATTICE_FCC };
    static Lattice sc(double /*a*/ = 1.0) { return Lattice(SC); }
    static Lattice bcc(double /*a*/ = 1.0) { return Lattice(BCC); }
    static Lattice fcc(double /*a*/ = 1.0) { return Lattice(FCC); }
    Lattice expand(uint32_t x, uint32_t y, uint32_t z) const {
        Lattice l = *this;
        l.lx_ *= x;
        l.ly_ *= y;
        l.lz_ *= z;
        return l;
    }
    Lattice expand_all(uint32_t n) const { return expand(n, n, n); }
    Lattice expand_x(uint32_t n) const { return expand(n, 1, 1); }
    Lattice expand_y(uint32_t n) const { return expand(1, n, 1); }
    Lattice expand_z(uint32_t n) const { return expand(1, 1, n); }
    Lattice drop_x() const { return drop(VG_PBC_X); }
    Lattice drop_y() const { return drop(VG_PBC_Y); }
    Lattice drop_z() const { return drop(VG_PBC_Z); }

    Cell cell() const { return cell_; }
    uint64_t num_sites() const {
        const uint64_t cells = (uint64_t)lx_ * ly_ * lz_;
        return cell_ == SC ? cells : (cell_ == BCC ? 2 * cells : 4 * cells);
    }
    vg_lattice descriptor() const {
        vg_lattice d;
        d.kind = cell_;
        d.lx = lx_;
        d.ly = ly_;
        d.lz = lz_;
        d.pbc = pbc_;
        return d;
    }
    // Neighbour slots of a site (sc: -x,+x,-y,+y,-z,+z; bcc: 8; fcc: 12, 4
    // per other sublattice), -1 when the neighbour would cross an open
    // boundary.
    std::vector<int64_t> neighbours(uint64_t site) const;

  private:
    explicit Lattice(Cell c) : cell_(c) {}
    Lattice drop(uint32_t bit) const {
        Lattice l = *this;
        l.pbc_ &= ~bit;
        return l;
    }
    Cell cell_;
    uint32_t lx_ = 1, ly_ = 1, lz_ = 1;
    uint32_t pbc_ = VG_PBC_X | VG_PBC_Y | VG_PBC_Z;
};

inline std::vector<int64_t> Lattice::neighbours(uint64_t site) const {
    auto step = [](int64_t a, int64_t d, int64_t n, bool periodic) -> int64_t {
        const int64_t b = a + d;
        if (b >= 0 && b < n) return b;
        if (!periodic) return -1;
        return ((b % n) + n) % n;
    };
    const bool px = pbc_ & VG_PBC_X, py = pbc_ & VG_PBC_Y, pz = pbc_ & VG_PBC_Z;
    std::vector<int64_t> out;
    if (cell_ == SC) {
        const int64_t x = site % lx_, y = (site / lx_) % ly_, z = site / ((uint64_t)lx_ * ly_);
        const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (auto& v : d) {
            const int64_t a = step(x, v[0], lx_, px), b = step(y, v[1], ly_, py),
                          c = step(z, v[2], lz_, pz);
            out.push_back(a < 0 || b < 0 || c < 0 ? -1 : a + (int64_t)lx_ * (b + (int64_t)ly_ * c));
        }
    } else if (cell_ == BCC) {
        const uint64_t cellid = site >> 1;
        const int sub = (int)(site & 1);
        const int64_t x = cellid % lx_, y = (cellid / lx_) % ly_,
                      z = cellid / ((uint64_t)lx_ * ly_);
        const int64_t sg = sub == 0 ? -1 : 1;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int64_t a = step(x, sg * dx, lx_, px), b = step(y, sg * dy, ly_, py),
                                  c = step(z, sg * dz, lz_, pz);
                    out.push_back(a < 0 || b < 0 || c < 0
                                      ? -1
                                      : 2 * (a + (int64_t)lx_ * (b + (int64_t)ly_ * c)) + (1 - sub));
                }
    } else {
        // Sublattice a bonds to 4 sites of each other sublattice b, in cells
        // c + d1*e1 + d2*e2 over the two axes where a and b differ, e = -1
        // where b sits higher in the cell than a, +1 where lower.
        static const int half[4][3] = {{0, 0, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
        const uint64_t cellid = site >> 2;
        const int a = (int)(site & 3);
        const int64_t c0[3] = {(int64_t)(cellid % lx_), (int64_t)((cellid / lx_) % ly_),
                               (int64_t)(cellid / ((uint64_t)lx_ * ly_))};
        const int64_t dims[3] = {lx_, ly_, lz_};
        const bool per[3] = {px, py, pz};
        for (int b = 0; b < 4; ++b) {
            if (b == a) continue;
            int ax[2], m = 0;
            for (int e = 0; e < 3; ++e)
                if (half[a][e] != half[b][e]) ax[m++] = e;
            for (int d2 = 0; d2 < 2; ++d2)
                for (int d1 = 0; d1 < 2; ++d1) {
                    int64_t q[3];
                    bool ok = true;
                    for (int e = 0; e < 3; ++e) {
                        int64_t d = (e == ax[0] ? d1 : 0) + (e == ax[1] ? d2 : 0);
                        if (half[b][e] > half[a][e]) d = -d;
                        q[e] = step(c0[e], d, dims[e], per[e]);
                        ok = ok && q[e] >= 0;
                    }
                    out.push_back(ok ? 4 * (q[0] + (int64_t)lx_ * (q[1] + (int64_t)ly_ * q[2])) + b
                                     : -1);
                }
        }
    }
    return out;
}

}  // namespace vegas
