// Host preparation of graph chains: a caller's exchange matrix (CSR) ->
// the colour classes and merged rows the graph passes read.
//
// The reference keeps Exchange::new's CsMat as given and evaluates row i of
// it for e_i (src/energy.rs:170-172, 194-206).  A parallel sweep needs sets
// of sites that share no nonzero, so the symmetrised pattern (j != i with
// J_ij or J_ji nonzero) is coloured greedily in site order: each site takes
// the smallest colour none of its lower-numbered neighbours holds.  Sites are
// then stored by (colour, id).  Site i's merged row lists its pattern
// neighbours by ascending id with a = the J_ij of row i (duplicates summed
// in entry order) and w = a + (the J_ji of row j, summed likewise) for a
// neighbour of lower colour, else 0; diagonal entries are summed, in entry
// order, into the constant the observables add.  The colouring, the order of
// every sum and the layout are a format the CPU restatement shares.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

namespace vgi {

int build_graph(const vg_csr& m, GraphBuild& out) {
    const uint64_t n = m.n_sites;
    if (n == 0 || n >= 0x7FFFFFFFull)
        return fail(VG_ERR_INVALID_ARGUMENT, "a graph chain needs 1 .. 2^31 - 2 sites (got %llu)",
                    (unsigned long long)n);
    if (!m.row_ptr || (!m.col && m.row_ptr[n]) || (!m.value && m.row_ptr[n]))
        return fail(VG_ERR_INVALID_ARGUMENT, "null CSR array");
    if (m.row_ptr[0] != 0) return fail(VG_ERR_INVALID_ARGUMENT, "row_ptr[0] must be 0");
    for (uint64_t i = 0; i < n; ++i)
        if (m.row_ptr[i + 1] < m.row_ptr[i])
            return fail(VG_ERR_INVALID_ARGUMENT, "row_ptr decreases at row %llu",
                        (unsigned long long)i);
    const uint64_t nnz = m.row_ptr[n];
    if (nnz >= 0x7FFFFFFFull) return fail(VG_ERR_INVALID_ARGUMENT, "too many nonzeros");
    struct Trip {
        uint64_t row, col, k;
        uint32_t kind;  // 0: entry (row, col) itself; 1: entry (col, row) seen from row
        double v;
    };
    std::vector<Trip> t;
    t.reserve(2 * nnz);
    out.diag = 0.0;
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t k = m.row_ptr[i]; k < m.row_ptr[i + 1]; ++k) {
            const uint64_t j = m.col[k];
            const double v = m.value[k];
            if (j >= n)
                return fail(VG_ERR_INVALID_ARGUMENT, "column %llu of row %llu is out of range",
                            (unsigned long long)j, (unsigned long long)i);
            if (!std::isfinite(v))
                return fail(VG_ERR_INVALID_ARGUMENT, "coupling of (%llu, %llu) is not finite",
                            (unsigned long long)i, (unsigned long long)j);
            if (j == i) {
                out.diag += v;
                continue;
            }
            t.push_back(Trip{i, j, k, 0u, v});
            t.push_back(Trip{j, i, k, 1u, v});
        }
    std::sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
        if (a.row != b.row) return a.row < b.row;
        if (a.col != b.col) return a.col < b.col;
        if (a.kind != b.kind) return a.kind < b.kind;
        return a.k < b.k;
    });
    // merged rows in site order: neighbour ids, a, and the transposed sums
    std::vector<uint64_t> mptr(n + 1, 0);
    std::vector<uint32_t> mcol;
    std::vector<double> ma, mb;
    mcol.reserve(t.size());
    for (size_t e = 0; e < t.size(); ++e) {
        if (e == 0 || t[e].row != t[e - 1].row || t[e].col != t[e - 1].col) {
            mcol.push_back((uint32_t)t[e].col);
            ma.push_back(0.0);
            mb.push_back(0.0);
            ++mptr[t[e].row + 1];
        }
        if (t[e].kind == 0)
            ma.back() = ma.back() + t[e].v;
        else
            mb.back() = mb.back() + t[e].v;
    }
    t.clear();
    t.shrink_to_fit();
    for (uint64_t i = 0; i < n; ++i) mptr[i + 1] += mptr[i];
    // greedy colouring in site order
    std::vector<uint32_t> colour(n, 0);
    uint64_t maxdeg = 0;
    for (uint64_t i = 0; i < n; ++i) maxdeg = std::max<uint64_t>(maxdeg, mptr[i + 1] - mptr[i]);
    std::vector<uint64_t> stamp(maxdeg + 2, 0);
    uint32_t ncol = 0;
    for (uint64_t i = 0; i < n; ++i) {
        for (uint64_t e = mptr[i]; e < mptr[i + 1]; ++e)
            if (mcol[e] < i) stamp[colour[mcol[e]]] = i + 1;
        uint32_t c = 0;
        while (stamp[c] == i + 1) ++c;
        colour[i] = c;
        ncol = std::max(ncol, c + 1);
    }
    // positions: sites sorted by (colour, id)
    out.ncol = ncol;
    out.cstart.assign(ncol + 1, 0);
    for (uint64_t i = 0; i < n; ++i) ++out.cstart[colour[i] + 1];
    for (uint32_t c = 0; c < ncol; ++c) out.cstart[c + 1] += out.cstart[c];
    out.max_class = 0;
    for (uint32_t c = 0; c < ncol; ++c)
        out.max_class = std::max(out.max_class, out.cstart[c + 1] - out.cstart[c]);
    out.site_of_pos.assign(n, 0);
    out.pos_of_site.assign(n, 0);
    std::vector<uint32_t> fill(out.cstart.begin(), out.cstart.end() - 1);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t p = fill[colour[i]]++;
        out.site_of_pos[p] = (uint32_t)i;
        out.pos_of_site[i] = p;
    }
    // merged rows in position order, neighbours as positions
    out.ptr.assign(n + 1, 0);
    out.nbr.clear();
    out.a.clear();
    out.w.clear();
    out.nbr.reserve(mcol.size());
    out.a.reserve(mcol.size());
    out.w.reserve(mcol.size());
    for (uint64_t p = 0; p < n; ++p) {
        const uint64_t i = out.site_of_pos[p];
        for (uint64_t e = mptr[i]; e < mptr[i + 1]; ++e) {
            const uint32_t j = mcol[e];
            out.nbr.push_back(out.pos_of_site[j]);
            out.a.push_back(ma[e]);
            out.w.push_back(colour[j] < colour[i] ? ma[e] + mb[e] : 0.0);
        }
        out.ptr[p + 1] = (uint32_t)out.nbr.size();
    }
    return VG_OK;
}

}  // namespace vgi
