// Lattice topology for the Hamiltonian's Exchange term: the unit cells and
// operations the reference takes from the vegas-lattice crate
// (src/input.rs:296-322, src/main.rs:25): Lattice::sc / bcc / fcc(a),
// expand(x,y,z), expand_all(n), expand_x/y/z(n), drop_x/y/z() (periodic ->
// open).  The engine needs a bipartite lattice: sc (site id x + lx*(y + ly*z))
// and bcc (site id 2*(cell) + sublattice); fcc is not two-colourable and is
// rejected when a chain is built from it.
#pragma once

#include <cstdint>
#include <vector>

#include "error.hpp"

namespace vegas {

class Lattice {
  public:
    enum Cell { SC = VG_LATTICE_SC, BCC = VG_LATTICE_BCC, FCC = 2 };
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
        if (cell_ == FCC)
            throw Error(ErrorKind::UnsupportedLattice,
                        "fcc is not bipartite; the two-colour sweep needs sc or bcc");
        vg_lattice d;
        d.kind = cell_;
        d.lx = lx_;
        d.ly = ly_;
        d.lz = lz_;
        d.pbc = pbc_;
        return d;
    }
    // Neighbour slots of a site (sc: -x,+x,-y,+y,-z,+z; bcc: 8), -1 when the
    // neighbour would cross an open boundary.
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
        throw Error(ErrorKind::UnsupportedLattice, "fcc neighbours are not provided");
    }
    return out;
}

}  // namespace vegas
