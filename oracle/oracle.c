/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h for who may use it).
 * Build: oracle/Makefile (gcc, -ffp-contract=off so every fp32 operation is
 * the single IEEE operation the checkerboard semantics name; the fused
 * multiply-adds they name are explicit fmaf calls).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11): 10 rounds of                        */
/*   (c0,c1,c2,c3) <- (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0))  */
/* with the key bumped by the Weyl constants between rounds.                 */
/* ------------------------------------------------------------------------ */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t a = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t b = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)b;
        uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)a;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static void draw(uint64_t seed, uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint32_t out[4]) {
    const uint32_t ctr[4] = {x, y, z, w};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox(ctr, key, out);
}

/* counter word w: colour | purpose << 1 | (sweep >> 32) << 8 */
/* counter word w: bit 0 colour & 1, bits 1-6 purpose, bit 7 colour >> 1,
 * bits 8-31 sweep >> 32 */
static uint32_t wword(uint64_t sweep, uint32_t purpose, uint32_t colour) {
    return (uint32_t)((sweep >> 32) << 8) | ((colour >> 1) << 7) | (purpose << 1) | (colour & 1u);
}

/* ------------------------------------------------------------------------ */
/* Lattice                                                                   */
/* ------------------------------------------------------------------------ */
uint64_t or_num_sites(const or_lattice* L) {
    uint64_t cells = (uint64_t)L->lx * L->ly * L->lz;
    return L->kind == 0 ? cells : (L->kind == 1 ? 2 * cells : 4 * cells);
}

int or_coordination(const or_lattice* L) { return L->kind == 0 ? 6 : (L->kind == 1 ? 8 : 12); }

/* colours of the checkerboard: sc and bcc are bipartite, fcc takes its four
 * sublattices */
static uint32_t ncolours(const or_lattice* L) { return L->kind == 2 ? 4u : 2u; }

/* fcc sublattice positions in half cells: corner, then the face centres
 * (0,1/2,1/2), (1/2,0,1/2), (1/2,1/2,0) */
static const int kFccSub[4][3] = {{0, 0, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

/* coordinate a+d along an axis of length n, or -1 when it leaves an open axis */
static int64_t step_axis(int64_t a, int64_t d, int64_t n, int periodic) {
    int64_t b = a + d;
    if (b >= 0 && b < n) return b;
    if (!periodic) return -1;
    return ((b % n) + n) % n;
}

void or_neighbours(const or_lattice* L, uint64_t site, int64_t* slots) {
    const int px = L->pbc & 1, py = (L->pbc >> 1) & 1, pz = (L->pbc >> 2) & 1;
    if (L->kind == 0) {
        int64_t x = (int64_t)(site % L->lx);
        int64_t y = (int64_t)((site / L->lx) % L->ly);
        int64_t z = (int64_t)(site / ((uint64_t)L->lx * L->ly));
        const int64_t d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int k = 0; k < 6; ++k) {
            int64_t a = step_axis(x, d[k][0], L->lx, px);
            int64_t b = step_axis(y, d[k][1], L->ly, py);
            int64_t c = step_axis(z, d[k][2], L->lz, pz);
            slots[k] = (a < 0 || b < 0 || c < 0) ? -1 : a + (int64_t)L->lx * (b + (int64_t)L->ly * c);
        }
    } else if (L->kind == 2) {
        /* fcc: a site of sublattice a in cell c bonds to 4 sites of every
         * other sublattice b, in cells c + d1*e1 + d2*e2 over the two axes
         * e1 < e2 where a and b differ, d1, d2 in {0,1}, with e = -1 along
         * an axis where b sits higher than a and +1 where lower.  Slots: b
         * ascending, then (d1, d2) = (0,0), (1,0), (0,1), (1,1). */
        uint64_t cell = site >> 2;
        int a = (int)(site & 3);
        int64_t xyz[3] = {(int64_t)(cell % L->lx), (int64_t)((cell / L->lx) % L->ly),
                          (int64_t)(cell / ((uint64_t)L->lx * L->ly))};
        const int64_t dims[3] = {L->lx, L->ly, L->lz};
        const int per[3] = {px, py, pz};
        int k = 0;
        for (int b = 0; b < 4; ++b) {
            if (b == a) continue;
            int ax[2], m = 0;
            for (int e = 0; e < 3; ++e)
                if (kFccSub[a][e] != kFccSub[b][e]) ax[m++] = e;
            for (int d2 = 0; d2 < 2; ++d2)
                for (int d1 = 0; d1 < 2; ++d1) {
                    int64_t q[3];
                    int ok = 1;
                    for (int e = 0; e < 3; ++e) {
                        int64_t d = 0;
                        if (e == ax[0]) d = d1;
                        if (e == ax[1]) d = d2;
                        if (kFccSub[b][e] > kFccSub[a][e]) d = -d;
                        q[e] = step_axis(xyz[e], d, dims[e], per[e]);
                        if (q[e] < 0) ok = 0;
                    }
                    slots[k++] = ok ? 4 * (q[0] + (int64_t)L->lx * (q[1] + (int64_t)L->ly * q[2])) + b
                                    : -1;
                }
        }
    } else {
        /* bcc: sublattice 0 (corner of cell x,y,z) neighbours the body centres
         * of cells (x-dx, y-dy, z-dz); sublattice 1 (centre of x,y,z) the
         * corners of cells (x+dx, y+dy, z+dz); dx,dy,dz in {0,1}. */
        uint64_t cell = site >> 1;
        int sub = (int)(site & 1);
        int64_t x = (int64_t)(cell % L->lx);
        int64_t y = (int64_t)((cell / L->lx) % L->ly);
        int64_t z = (int64_t)(cell / ((uint64_t)L->lx * L->ly));
        int64_t sgn = sub == 0 ? -1 : 1;
        int k = 0;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    int64_t a = step_axis(x, sgn * dx, L->lx, px);
                    int64_t b = step_axis(y, sgn * dy, L->ly, py);
                    int64_t c = step_axis(z, sgn * dz, L->lz, pz);
                    slots[k++] = (a < 0 || b < 0 || c < 0)
                                     ? -1
                                     : 2 * (a + (int64_t)L->lx * (b + (int64_t)L->ly * c)) + (1 - sub);
                }
    }
}

void or_place(const or_lattice* L, uint64_t site, uint32_t* colour, uint32_t* row, uint32_t* pos) {
    if (L->kind == 0) {
        uint32_t x = (uint32_t)(site % L->lx);
        uint32_t y = (uint32_t)((site / L->lx) % L->ly);
        uint32_t z = (uint32_t)(site / ((uint64_t)L->lx * L->ly));
        *colour = (x + y + z) & 1u;
        *row = y + L->ly * z;
        *pos = x / 2;
    } else {
        const int sh = L->kind == 2 ? 2 : 1;
        uint64_t cell = site >> sh;
        uint32_t x = (uint32_t)(cell % L->lx);
        uint32_t y = (uint32_t)((cell / L->lx) % L->ly);
        uint32_t z = (uint32_t)(cell / ((uint64_t)L->lx * L->ly));
        *colour = (uint32_t)(site & ((1u << sh) - 1u));
        *row = y + L->ly * z;
        *pos = x;
    }
}

/* natural site of (colour, row, position) */
static uint64_t site_of(const or_lattice* L, uint32_t colour, uint32_t row, uint32_t pos) {
    uint32_t y = row % L->ly, z = row / L->ly;
    if (L->kind == 0) {
        uint32_t x = 2 * pos + ((colour + y + z) & 1u);
        return x + (uint64_t)L->lx * (y + (uint64_t)L->ly * z);
    }
    return ncolours(L) * (pos + (uint64_t)L->lx * (y + (uint64_t)L->ly * z)) + colour;
}

static uint32_t row_length(const or_lattice* L) { return L->kind == 0 ? L->lx / 2 : L->lx; }

/* ------------------------------------------------------------------------ */
/* Ising checkerboard                                                        */
/* ------------------------------------------------------------------------ */
void or_ising_table(const or_lattice* L, const or_model* M, uint32_t* thr, uint32_t* always) {
    const int z = or_coordination(L);
    const double J = (M->terms & 1u) ? M->exchange : 0.0;
    const double hm = (M->terms & 2u) ? fabs(M->field) : 0.0;
    const double hz = M->field_dir[2] >= 0.0 ? 1.0 : -1.0;
    const double T = M->temperature < DBL_EPSILON ? DBL_EPSILON : M->temperature;
    for (int b = 0; b < 2; ++b) {
        always[b] = 0;
        double s = b ? 1.0 : -1.0;
        for (int S = 0; S < OR_THR_STRIDE; ++S) {
            thr[b * OR_THR_STRIDE + S] = 0;
            if (S > 2 * z) continue;
            double sigma = (double)(S - z);
            /* flip s -> -s: exchange term 2 J s sigma, Zeeman term -2 |h| hz s */
            double dE = 2.0 * J * s * sigma - 2.0 * hm * hz * s;
            if (dE <= 0.0) {
                always[b] |= 1u << S;
                continue;
            }
            double p = exp(-dE / T);
            double th = ceil(p * 4294967296.0);
            if (th >= 4294967296.0)
                always[b] |= 1u << S;
            else
                thr[b * OR_THR_STRIDE + S] = (uint32_t)th;
        }
    }
}

void or_ising_init(const or_lattice* L, uint64_t seed, int8_t* spins) {
    uint64_t n = or_num_sites(L);
    for (uint64_t site = 0; site < n; ++site) {
        uint32_t c, row, pos, r[4];
        or_place(L, site, &c, &row, &pos);
        draw(seed, pos >> 2, row, 0, wword(0, 3, c), r);
        spins[site] = (r[pos & 3] >> 31) == 0 ? 1 : -1;
    }
}

/* High byte of the acceptance draw of position pos (purpose IsingA).
 * bcc and fcc: byte pos & 3 of word (pos & 15) >> 2 of the call keyed by
 * pos >> 4.  sc (the multispin pass): the 32 positions of a 32-aligned
 * segment share the calls keyed by 2 * (pos >> 5) and 2 * (pos >> 5) + 1;
 * bit k of the byte is bit pos & 31 of word k & 3 of the first call
 * (k < 4) or of the second (k >= 4). */
static uint32_t bcc_fcc_high_byte(uint64_t seed, uint32_t pos, uint32_t row, uint64_t t,
                                  uint32_t c) {
    uint32_t ra[4];
    draw(seed, pos >> 4, row, (uint32_t)t, wword(t, 0, c), ra);
    return (ra[(pos & 15) >> 2] >> (8 * (pos & 3))) & 0xFFu;
}

static uint32_t sc_high_byte(uint64_t seed, uint32_t pos, uint32_t row, uint64_t t, uint32_t c) {
    uint32_t r0[4], r1[4], v = 0;
    draw(seed, 2 * (pos >> 5), row, (uint32_t)t, wword(t, 0, c), r0);
    draw(seed, 2 * (pos >> 5) + 1, row, (uint32_t)t, wword(t, 0, c), r1);
    for (int k = 0; k < 4; ++k) {
        v |= ((r0[k] >> (pos & 31)) & 1u) << k;
        v |= ((r1[k] >> (pos & 31)) & 1u) << (k + 4);
    }
    return v;
}

void or_ising_sweeps(const or_lattice* L, const or_model* M, uint64_t seed, uint64_t sweep0,
                     uint64_t nsweeps, int8_t* spins, int64_t* sums) {
    const uint64_t n = or_num_sites(L);
    const int z = or_coordination(L);
    uint32_t thr[2 * OR_THR_STRIDE], always[2];
    or_ising_table(L, M, thr, always);
    int64_t nb[12];
    for (uint64_t k = 0; k < nsweeps; ++k) {
        const uint64_t t = sweep0 + k;
        for (uint32_t c = 0; c < ncolours(L); ++c) {
            for (uint64_t site = 0; site < n; ++site) {
                uint32_t cc, row, pos;
                or_place(L, site, &cc, &row, &pos);
                if (cc != c) continue;
                or_neighbours(L, site, nb);
                uint32_t S = 0;
                for (int j = 0; j < z; ++j) S += nb[j] < 0 ? 1u : (spins[nb[j]] > 0 ? 2u : 0u);
                uint32_t b = spins[site] > 0;
                if (M->proposal == 1) {
                    uint32_t r[4];
                    draw(seed, pos >> 7, row, (uint32_t)t, wword(t, 2, c), r);
                    uint32_t bit = (r[(pos >> 5) & 3] >> (pos & 31)) & 1u;
                    if (bit != b) continue; /* proposal equals the current spin */
                }
                int flip;
                if ((always[b] >> S) & 1u) {
                    flip = 1;
                } else {
                    uint32_t rb[4];
                    draw(seed, pos >> 2, row, (uint32_t)t, wword(t, 1, c), rb);
                    uint32_t hi8 = L->kind == 0 ? sc_high_byte(seed, pos, row, t, c)
                                                : bcc_fcc_high_byte(seed, pos, row, t, c);
                    uint32_t lo24 = rb[pos & 3] >> 8;
                    uint32_t draw32 = (hi8 << 24) | lo24;
                    flip = draw32 < thr[b * OR_THR_STRIDE + S];
                }
                if (flip) spins[site] = (int8_t)-spins[site];
            }
        }
        if (sums) {
            int64_t bond = 0, mag = 0;
            /* every bond once: from each site to its lower-colour neighbours */
            for (uint64_t site = 0; site < n; ++site) {
                mag += spins[site];
                uint32_t cc, row, pos;
                or_place(L, site, &cc, &row, &pos);
                if (cc == 0) continue;
                or_neighbours(L, site, nb);
                for (int j = 0; j < z; ++j) {
                    if (nb[j] < 0) continue;
                    uint32_t cn, rn, pn;
                    or_place(L, (uint64_t)nb[j], &cn, &rn, &pn);
                    if (cn < cc) bond += (int64_t)spins[site] * spins[nb[j]];
                }
            }
            int64_t* o = sums + 5 * k;
            o[0] = bond;
            o[1] = 0;
            o[2] = 0;
            o[3] = mag;
            o[4] = (int64_t)n;
        }
    }
}

/* The same sc checkerboard, written for speed rather than generality: the
 * CPU baseline's best-CPU line (two Philox calls per 32 positions for the
 * high bytes, as the GPU does; rows of a colour shared among `threads`
 * OpenMP threads).  Every draw is the one or_ising_sweeps makes, so the
 * result equals it bit for bit for any thread count. */
void or_ising_sweeps_sc_fast(const or_lattice* L, const or_model* M, uint64_t seed,
                             uint64_t sweep0, uint64_t nsweeps, int8_t* spins, int64_t* sums,
                             int threads) {
    const uint32_t lx = L->lx, ly = L->ly, lz = L->lz, w = lx / 2, rows = ly * lz;
    const int px = L->pbc & 1, py = (L->pbc >> 1) & 1, pz = (L->pbc >> 2) & 1;
    uint32_t thr[2 * OR_THR_STRIDE], always[2];
    or_ising_table(L, M, thr, always);
    const int64_t sx = 1, sy = lx, sz = (int64_t)lx * ly;
    for (uint64_t k = 0; k < nsweeps; ++k) {
        const uint64_t t = sweep0 + k;
        for (uint32_t c = 0; c < 2; ++c) {
#pragma omp parallel for schedule(static) num_threads(threads)
            for (uint32_t row = 0; row < rows; ++row) {
                const uint32_t y = row % ly, zz = row / ly;
                const int64_t base = sy * y + sz * zz;
                /* neighbour offsets along y and z, 0 when absent (open) */
                const int ym = y > 0 || py, yp = y + 1 < ly || py;
                const int zm = zz > 0 || pz, zp = zz + 1 < lz || pz;
                const int64_t dym = y > 0 ? -sy : sy * (ly - 1), dyp = y + 1 < ly ? sy : -sy * (ly - 1);
                const int64_t dzm = zz > 0 ? -sz : sz * (lz - 1), dzp = zz + 1 < lz ? sz : -sz * (lz - 1);
                const uint32_t off = (c + y + zz) & 1u;
                uint32_t ra[8] = {0, 0, 0, 0, 0, 0, 0, 0}, rp[4] = {0, 0, 0, 0};
                for (uint32_t pos = 0; pos < w; ++pos) {
                    if ((pos & 31) == 0) {
                        draw(seed, 2 * (pos >> 5), row, (uint32_t)t, wword(t, 0, c), ra);
                        draw(seed, 2 * (pos >> 5) + 1, row, (uint32_t)t, wword(t, 0, c), ra + 4);
                    }
                    if (M->proposal == 1 && (pos & 127) == 0)
                        draw(seed, pos >> 7, row, (uint32_t)t, wword(t, 2, c), rp);
                    const uint32_t x = 2 * pos + off;
                    const int64_t site = base + x;
                    const int xm = x > 0 || px, xp = x + 1 < lx || px;
                    const int64_t dxm = x > 0 ? -sx : sx * (lx - 1), dxp = x + 1 < lx ? sx : -sx * (lx - 1);
                    uint32_t S = 0;
                    S += xm ? (spins[site + dxm] > 0 ? 2u : 0u) : 1u;
                    S += xp ? (spins[site + dxp] > 0 ? 2u : 0u) : 1u;
                    S += ym ? (spins[site + dym] > 0 ? 2u : 0u) : 1u;
                    S += yp ? (spins[site + dyp] > 0 ? 2u : 0u) : 1u;
                    S += zm ? (spins[site + dzm] > 0 ? 2u : 0u) : 1u;
                    S += zp ? (spins[site + dzp] > 0 ? 2u : 0u) : 1u;
                    const uint32_t b = spins[site] > 0;
                    if (M->proposal == 1 && ((rp[(pos >> 5) & 3] >> (pos & 31)) & 1u) != b) continue;
                    int flip;
                    if ((always[b] >> S) & 1u) {
                        flip = 1;
                    } else {
                        const uint32_t th = thr[b * OR_THR_STRIDE + S];
                        uint32_t hi8 = 0;
                        for (int kb = 0; kb < 8; ++kb) hi8 |= ((ra[kb] >> (pos & 31)) & 1u) << kb;
                        if (hi8 != th >> 24) {
                            flip = hi8 < th >> 24;
                        } else {
                            uint32_t rb[4];
                            draw(seed, pos >> 2, row, (uint32_t)t, wword(t, 1, c), rb);
                            flip = (rb[pos & 3] >> 8) < (th & 0xFFFFFFu);
                        }
                    }
                    if (flip) spins[site] = (int8_t)-spins[site];
                }
            }
        }
        if (sums) {
            int64_t bond = 0, mag = 0;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : bond, mag)
            for (uint32_t row = 0; row < rows; ++row) {
                const uint32_t y = row % ly, zz = row / ly;
                const int64_t base = sy * y + sz * zz;
                for (uint32_t x = 0; x < lx; ++x) {
                    const int64_t site = base + x;
                    mag += spins[site];
                    if (((x + y + zz) & 1u) != 1u) continue;  /* bonds from colour-1 sites */
                    int64_t nbsum = 0;
                    if (x > 0 || px) nbsum += spins[x > 0 ? site - sx : site + sx * (lx - 1)];
                    if (x + 1 < lx || px) nbsum += spins[x + 1 < lx ? site + sx : site - sx * (lx - 1)];
                    if (y > 0 || py) nbsum += spins[y > 0 ? site - sy : site + sy * (ly - 1)];
                    if (y + 1 < ly || py) nbsum += spins[y + 1 < ly ? site + sy : site - sy * (ly - 1)];
                    if (zz > 0 || pz) nbsum += spins[zz > 0 ? site - sz : site + sz * (lz - 1)];
                    if (zz + 1 < lz || pz) nbsum += spins[zz + 1 < lz ? site + sz : site - sz * (lz - 1)];
                    bond += spins[site] * nbsum;
                }
            }
            int64_t* o = sums + 5 * k;
            o[0] = bond;
            o[1] = 0;
            o[2] = 0;
            o[3] = mag;
            o[4] = (int64_t)lx * ly * lz;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Heisenberg checkerboard (fp32; every line is one IEEE operation or fmaf) */
/* ------------------------------------------------------------------------ */
float or_exp_nonpos(float x) {
    if (x < -86.0f) return 0.0f;
    const float log2e = (float)1.4426950408889634;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = (float)(0.6931471805599453 - 0.693145751953125);
    float n = rintf(x * log2e);
    float r = fmaf(n, -ln2_hi, x);
    r = fmaf(n, -ln2_lo, r);
    const float c[8] = {1.0f, 1.0f, 0.5f, (float)(1.0 / 6.0), (float)(1.0 / 24.0),
                        (float)(1.0 / 120.0), (float)(1.0 / 720.0), (float)(1.0 / 5040.0)};
    float p = c[7];
    for (int k = 6; k >= 0; --k) p = fmaf(p, r, c[k]);
    return ldexpf(p, (int)n);
}

/* a . b as one product and two fused multiply-adds, x then y then z. */
static float dot3f(const float* a, const float* b) {
    const float t = a[0] * b[0];
    return fmaf(a[2], b[2], fmaf(a[1], b[1], t));
}

static float unit24f(uint32_t w) { return (float)(w >> 8) * 5.9604644775390625e-08f; }
static float sym24f(uint32_t w) { return (float)(w >> 8) * 1.1920928955078125e-07f - 1.0f; }

static int32_t fix27(float v) { return (int32_t)lrintf(v * 134217728.0f); }

/* Marsaglia (1972) in fp32 from a pair of words; returns 0 when rejected. */
static int marsaglia_f32(uint32_t w1, uint32_t w2, float* out) {
    float x1 = sym24f(w1), x2 = sym24f(w2);
    float a = x1 * x1;
    float b = x2 * x2;
    float q = a + b;
    if (!(q < 1.0f)) return 0;
    float root = sqrtf(1.0f - q);
    out[0] = (2.0f * x1) * root;
    out[1] = (2.0f * x2) * root;
    out[2] = 1.0f - 2.0f * q;
    return 1;
}

void or_heis_init(const or_lattice* L, uint64_t seed, float* xyz) {
    uint64_t n = or_num_sites(L);
    for (uint64_t site = 0; site < n; ++site) {
        uint32_t c, row, pos;
        or_place(L, site, &c, &row, &pos);
        float s[3] = {0.0f, 0.0f, 1.0f};
        int done = 0;
        for (uint32_t k = 0; k < 64 && !done; ++k) {
            uint32_t r[4];
            draw(seed, pos, row, k, wword(0, 4, c), r);
            for (int t = 0; t < 2 && !done; ++t) done = marsaglia_f32(r[2 * t], r[2 * t + 1], s);
        }
        memcpy(xyz + 3 * site, s, sizeof s);
    }
}

/* 24-bit value m (0..4) of a Philox call: bits [24m, 24m+24) of the 128-bit
 * string w0 | w1 << 32 | w2 << 64 | w3 << 96 (bits 120-127 unused). */
static uint32_t v24(const uint32_t w[4], int m) {
    const uint64_t lo = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    const uint64_t hi = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
    uint64_t v;
    switch (m) {
        case 0: v = lo; break;
        case 1: v = lo >> 24; break;
        case 2: v = (lo >> 48) | (hi << 16); break;
        case 3: v = hi >> 8; break;
        default: v = hi >> 32; break;
    }
    return (uint32_t)(v & 0xFFFFFFu);
}

static float f24u(uint32_t v) { return (float)v * 5.9604644775390625e-08f; }
static float f24s(uint32_t v) { return (float)v * 1.1920928955078125e-07f - 1.0f; }

/* Marsaglia try from two 24-bit values; 0 when the point misses the disc. */
static int marsaglia24(uint32_t a, uint32_t b, float* out) {
    float x1 = f24s(a), x2 = f24s(b);
    float p = x1 * x1;
    float q2 = x2 * x2;
    float q = p + q2;
    if (!(q < 1.0f)) return 0;
    float root = sqrtf(1.0f - q);
    out[0] = (2.0f * x1) * root;
    out[1] = (2.0f * x2) * root;
    out[2] = 1.0f - 2.0f * q;
    return 1;
}

/* Fresh uniform-sphere proposal of one site: Philox calls k = 0, 1, ...
 * with counter (pos, row, sweep, purpose 16 + k); u = value 0 of call 0;
 * tries (1,2), (3,4) of call 0, (0,1), (2,3) of call 1, then (0,1), (2,3)
 * of calls 2..15; exhausted -> (0,0,1). */
static float random_proposal(uint64_t seed, uint32_t pos, uint32_t row, uint64_t t, uint32_t c,
                             float* p) {
    uint32_t w0[4], w1[4];
    draw(seed, pos, row, (uint32_t)t, wword(t, 16, c), w0);
    draw(seed, pos, row, (uint32_t)t, wword(t, 17, c), w1);
    const float u = f24u(v24(w0, 0));
    if (marsaglia24(v24(w0, 1), v24(w0, 2), p)) return u;
    if (marsaglia24(v24(w0, 3), v24(w0, 4), p)) return u;
    if (marsaglia24(v24(w1, 0), v24(w1, 1), p)) return u;
    if (marsaglia24(v24(w1, 2), v24(w1, 3), p)) return u;
    for (uint32_t k = 2; k < 16; ++k) {
        uint32_t w[4];
        draw(seed, pos, row, (uint32_t)t, wword(t, 16 + k, c), w);
        if (marsaglia24(v24(w, 0), v24(w, 1), p)) return u;
        if (marsaglia24(v24(w, 2), v24(w, 3), p)) return u;
    }
    p[0] = 0.0f;
    p[1] = 0.0f;
    p[2] = 1.0f;
    return u;
}

void or_heis_sweeps(const or_lattice* L, const or_model* M, uint64_t seed, uint64_t sweep0,
                    uint64_t nsweeps, float* xyz, int64_t* sums) {
    const int z = or_coordination(L);
    const uint32_t rows = L->ly * L->lz, w = row_length(L);
    const float jn = (M->terms & 1u) ? (float)(-M->exchange) : 0.0f;
    const float kk = (M->terms & 4u) ? (float)M->anisotropy : 0.0f;
    const float hm = (M->terms & 2u) ? (float)fabs(M->field) : 0.0f;
    const double T = M->temperature < DBL_EPSILON ? DBL_EPSILON : M->temperature;
    const float beta = (float)(1.0 / T);
    const float axis[3] = {(float)M->axis[0], (float)M->axis[1], (float)M->axis[2]};
    const float fdir[3] = {(float)M->field_dir[0], (float)M->field_dir[1], (float)M->field_dir[2]};
    int64_t nb[12];
    for (uint64_t k = 0; k < nsweeps; ++k) {
        const uint64_t t = sweep0 + k;
        int64_t acc[5] = {0, 0, 0, 0, 0};
        for (uint32_t c = 0; c < ncolours(L); ++c) {
            for (uint32_t row = 0; row < rows; ++row) {
                for (uint32_t g = 0; 4 * g < w; ++g) {
                    /* flip proposals: word (pos & 3) of one call per 4 positions */
                    uint32_t fw[4];
                    if (M->proposal != 1) draw(seed, g, row, (uint32_t)t, wword(t, 16, c), fw);
                    for (uint32_t pos = 4 * g; pos < 4 * g + 4 && pos < w; ++pos) {
                        const uint64_t site = site_of(L, c, row, pos);
                        float* s = xyz + 3 * site;
                        or_neighbours(L, site, nb);
                        /* h: slots summed left to right; hb: the same sum
                         * over the slots of lower colours (the first
                         * z/ncol*c slots on fcc; every slot of a colour-1
                         * site on sc and bcc), for the bond sum */
                        const int nlow = L->kind == 2 ? 4 * (int)c : (c == 1 ? z : 0);
                        float h[3], hb[3] = {0.0f, 0.0f, 0.0f};
                        for (int a = 0; a < 3; ++a) {
                            float v0 = nb[0] < 0 ? 0.0f : xyz[3 * nb[0] + a];
                            float v1 = nb[1] < 0 ? 0.0f : xyz[3 * nb[1] + a];
                            float acc_h = v0 + v1;
                            for (int j = 2; j < z; ++j) {
                                if (j == nlow) hb[a] = acc_h;
                                acc_h = acc_h + (nb[j] < 0 ? 0.0f : xyz[3 * nb[j] + a]);
                            }
                            if (nlow == z) hb[a] = acc_h;
                            h[a] = acc_h;
                        }
                        float p[3] = {-s[0], -s[1], -s[2]};
                        float u;
                        if (M->proposal == 1)
                            u = random_proposal(seed, pos, row, t, c, p);
                        else
                            u = unit24f(fw[pos & 3]);
                        const float d[3] = {p[0] - s[0], p[1] - s[1], p[2] - s[2]};
                        const float e_ex = jn * dot3f(d, h);
                        const float pe = dot3f(p, axis), oe = dot3f(s, axis);
                        const float oe2 = oe * oe;
                        const float e_an = kk * fmaf(pe, pe, -oe2);
                        const float e_ze = hm * dot3f(d, fdir);
                        const float e1 = e_ex + e_an;
                        const float dE = e1 + e_ze;
                        int accept = dE < 0.0f;
                        if (!accept) {
                            const float x = -(dE * beta);
                            accept = u < or_exp_nonpos(x);
                        }
                        if (accept) memcpy(s, p, sizeof p);
                        acc[1] += fix27(s[0]);
                        acc[2] += fix27(s[1]);
                        acc[3] += fix27(s[2]);
                        const float se = dot3f(s, axis);
                        acc[4] += fix27(se * se);
                        if (nlow > 0) acc[0] += fix27(dot3f(s, hb));
                    }
                }
            }
        }
        if (sums) memcpy(sums + 5 * k, acc, sizeof acc);
    }
}

/* ------------------------------------------------------------------------ */
/* Wolff single-cluster steps (integrator.rs:146-230) with the engine's     */
/* draws: the seed site is floor(r64 * N / 2^64) for r64 = words 0,1 of     */
/* Philox(0, 0, step, w(step, 6, 0)); a bond between two equal spins of the */
/* cluster's sign is active iff word 0 of Philox(lo, slot, step, w(step, 5, */
/* 0)) is below ceil((1 - exp(-2J/T)) * 2^32), (lo, slot) being the bond's  */
/* endpoint with the lower site id and its neighbour slot.  The cluster is  */
/* the seed's connected component over active bonds, so it does not depend */
/* on the order the search visits it in.                                    */
/* ------------------------------------------------------------------------ */
/* slot of neighbour j's list that holds the bond of slot k of site i */
static int back_slot(const or_lattice* L, uint64_t i, int k) {
    if (L->kind == 0) return k ^ 1; /* -x <-> +x, -y <-> +y, -z <-> +z */
    if (L->kind == 1) return k;     /* (dz, dy, dx) from either sublattice */
    const int a = (int)(i & 3), r = k >> 2, b = r + (r >= a ? 1 : 0);
    return 4 * (a - (a > b ? 1 : 0)) + (k & 3);
}

void or_wolff_threshold(const or_model* M, uint32_t* thr, int* all) {
    const double T = M->temperature < DBL_EPSILON ? DBL_EPSILON : M->temperature;
    const double J = (M->terms & 1u) ? M->exchange : 0.0;
    const double p = 1.0 - exp(-2.0 * J / T);
    *thr = 0;
    *all = 0;
    if (!(p > 0.0)) return;
    const double th = ceil(p * 4294967296.0);
    if (th >= 4294967296.0)
        *all = 1;
    else
        *thr = (uint32_t)th;
}

void or_wolff_steps(const or_lattice* L, const or_model* M, uint64_t seed, uint64_t step0,
                    uint64_t nsteps, int8_t* spins, int64_t* sums) {
    const uint64_t n = or_num_sites(L);
    const int z = or_coordination(L);
    uint32_t thr;
    int all;
    or_wolff_threshold(M, &thr, &all);
    uint64_t* queue = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint8_t* in = (uint8_t*)malloc(n);
    int64_t nb[12];
    for (uint64_t k = 0; k < nsteps; ++k) {
        const uint64_t t = step0 + k;
        uint32_t r[4];
        draw(seed, 0, 0, (uint32_t)t, wword(t, 6, 0), r);
        const uint64_t r64 = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
        const uint64_t src = (uint64_t)(((unsigned __int128)r64 * n) >> 64);
        const int8_t ref = spins[src];
        memset(in, 0, n);
        in[src] = 1;
        queue[0] = src;
        uint64_t head = 0, tail = 1;
        while (head < tail) {
            const uint64_t i = queue[head++];
            or_neighbours(L, i, nb);
            for (int sl = 0; sl < z; ++sl) {
                const int64_t j = nb[sl];
                if (j < 0 || (uint64_t)j == i || in[j] || spins[j] != ref) continue;
                uint64_t lo;
                int ls;
                if (i < (uint64_t)j) {
                    lo = i;
                    ls = sl;
                } else {
                    lo = (uint64_t)j;
                    ls = back_slot(L, i, sl);
                }
                uint32_t b[4];
                draw(seed, (uint32_t)lo, (uint32_t)ls, (uint32_t)t, wword(t, 5, 0), b);
                if (all || b[0] < thr) {
                    in[j] = 1;
                    queue[tail++] = (uint64_t)j;
                }
            }
        }
        for (uint64_t q = 0; q < tail; ++q) spins[queue[q]] = (int8_t)-ref;
        if (sums) {
            int64_t bond = 0, mag = 0;
            for (uint64_t i = 0; i < n; ++i) {
                mag += spins[i];
                or_neighbours(L, i, nb);
                for (int sl = 0; sl < z; ++sl)
                    if (nb[sl] > (int64_t)i) bond += (int64_t)spins[i] * spins[nb[sl]];
            }
            int64_t* o = sums + 5 * k;
            o[0] = bond;
            o[1] = 0;
            o[2] = 0;
            o[3] = mag;
            o[4] = (int64_t)n;
        }
    }
    free(queue);
    free(in);
}

/* ------------------------------------------------------------------------ */
/* PCG64: 128-bit LCG with multiplier 0x2360ED051FC65DA44385DF649FCCF645,   */
/* output XSL-RR of the stepped state (O'Neill, PCG, 2014).                  */
/* ------------------------------------------------------------------------ */
typedef unsigned __int128 u128;

static u128 pcg_mult(void) {
    return ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;
}

uint64_t or_pcg64_next(or_pcg64* g) {
    u128 state = ((u128)g->state_hi << 64) | g->state_lo;
    u128 inc = ((u128)g->inc_hi << 64) | g->inc_lo;
    state = state * pcg_mult() + inc;
    g->state_hi = (uint64_t)(state >> 64);
    g->state_lo = (uint64_t)state;
    uint64_t xsl = (uint64_t)(state >> 64) ^ (uint64_t)state;
    unsigned rot = (unsigned)(state >> 122);
    return (xsl >> rot) | (xsl << ((64 - rot) & 63));
}

static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_pcg64_seed(or_pcg64* g, uint64_t seed) {
    uint64_t x = seed;
    g->state_hi = splitmix64(&x);
    g->state_lo = splitmix64(&x);
    g->inc_hi = splitmix64(&x);
    g->inc_lo = splitmix64(&x) | 1u;
}

static double pcg_f64(or_pcg64* g) { return (double)(or_pcg64_next(g) >> 11) * 0x1.0p-53; }

/* unbiased integer in [0, n) by Lemire's multiply-and-reject method */
static uint64_t pcg_below(or_pcg64* g, uint64_t n) {
    u128 m = (u128)or_pcg64_next(g) * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
        uint64_t t = (0 - n) % n;
        while (l < t) {
            m = (u128)or_pcg64_next(g) * n;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

/* ------------------------------------------------------------------------ */
/* Sequential reference restatement                                         */
/* ------------------------------------------------------------------------ */
struct or_seq {
    or_lattice L;
    or_model M;
    uint64_t n;
    uint64_t* indptr;  /* CSR of the exchange matrix: usize indices, f64 values */
    uint64_t* indices;
    double* values;
    int8_t* ising;
    double* heis;
    or_pcg64 rng;
};

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

or_seq* or_seq_create(const or_lattice* L, const or_model* M, uint64_t seed) {
    or_seq* s = (or_seq*)calloc(1, sizeof(or_seq));
    s->L = *L;
    s->M = *M;
    s->n = or_num_sites(L);
    const int z = or_coordination(L);
    /* Exchange::from_lattice: J on every bond, both directions; duplicate
     * entries (length-2 periodic axes) summed, columns ascending. */
    s->indptr = (uint64_t*)malloc((s->n + 1) * sizeof(uint64_t));
    s->indices = (uint64_t*)malloc(s->n * z * sizeof(uint64_t));
    s->values = (double*)malloc(s->n * z * sizeof(double));
    uint64_t nnz = 0;
    int64_t nb[12];
    for (uint64_t i = 0; i < s->n; ++i) {
        s->indptr[i] = nnz;
        or_neighbours(L, i, nb);
        int m = 0;
        for (int j = 0; j < z; ++j)
            if (nb[j] >= 0) nb[m++] = nb[j];
        qsort(nb, (size_t)m, sizeof(int64_t), cmp_i64);
        for (int j = 0; j < m; ++j) {
            if (j > 0 && nb[j] == nb[j - 1]) {
                s->values[nnz - 1] += M->exchange;
            } else {
                s->indices[nnz] = (uint64_t)nb[j];
                s->values[nnz] = M->exchange;
                ++nnz;
            }
        }
    }
    s->indptr[s->n] = nnz;
    if (M->spin == 2)
        s->heis = (double*)calloc(3 * s->n, sizeof(double));
    else
        s->ising = (int8_t*)calloc(s->n, 1);
    or_pcg64_seed(&s->rng, seed);
    return s;
}

void or_seq_destroy(or_seq* s) {
    if (!s) return;
    free(s->indptr);
    free(s->indices);
    free(s->values);
    free(s->ising);
    free(s->heis);
    free(s);
}

uint64_t or_seq_nnz(const or_seq* s) { return s->indptr[s->n]; }

void or_seq_set_thermostat(or_seq* s, double temperature, double field, const double dir[3]) {
    s->M.temperature = temperature < DBL_EPSILON ? DBL_EPSILON : temperature;
    s->M.field = field;
    for (int k = 0; k < 3; ++k) s->M.field_dir[k] = dir ? dir[k] : (k == 2 ? 1.0 : 0.0);
}

void or_seq_set_state(or_seq* s, const void* spins) {
    if (s->ising)
        memcpy(s->ising, spins, s->n);
    else
        memcpy(s->heis, spins, 3 * s->n * sizeof(double));
}

void or_seq_get_state(const or_seq* s, void* spins) {
    if (s->ising)
        memcpy(spins, s->ising, s->n);
    else
        memcpy(spins, s->heis, 3 * s->n * sizeof(double));
}

/* s_i . v for a spin given as three components */
static double sdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void spin_vec(const or_seq* s, uint64_t i, double* v) {
    if (s->ising) {
        v[0] = 0.0;
        v[1] = 0.0;
        v[2] = (double)s->ising[i];
    } else {
        v[0] = s->heis[3 * i];
        v[1] = s->heis[3 * i + 1];
        v[2] = s->heis[3 * i + 2];
    }
}

/* Ising dot products are +-1 (state.rs:95-101); the field and anisotropy
 * reference of an Ising system is Up or Down. */
static void ref_vec(const or_seq* s, const double* d, double* v) {
    if (s->ising) {
        v[0] = 0.0;
        v[1] = 0.0;
        v[2] = d[2] >= 0.0 ? 1.0 : -1.0;
    } else {
        v[0] = d[0];
        v[1] = d[1];
        v[2] = d[2];
    }
}

static double e_exchange(const or_seq* s, uint64_t i) {
    double si[3], sj[3], e = 0.0;
    spin_vec(s, i, si);
    for (uint64_t k = s->indptr[i]; k < s->indptr[i + 1]; ++k) {
        spin_vec(s, s->indices[k], sj);
        e += -s->values[k] * sdot(si, sj);
    }
    return e;
}

/* Compound(Compound(Compound(Exchange, Zeeman), Anisotropy), Gauge) energy of
 * site i with the absent terms left out (energy.rs:254-256). */
static double site_energy(const or_seq* s, uint64_t i) {
    const or_model* M = &s->M;
    double si[3], r[3];
    spin_vec(s, i, si);
    double e = 0.0;
    int first = 1;
    if (M->terms & 1u) {
        e = e_exchange(s, i);
        first = 0;
    }
    if (M->terms & 2u) {
        ref_vec(s, M->field_dir, r);
        double z = sdot(si, r) * fabs(M->field);
        e = first ? z : e + z;
        first = 0;
    }
    if (M->terms & 4u) {
        ref_vec(s, M->axis, r);
        double d = sdot(si, r);
        double a = d * d * M->anisotropy;
        e = first ? a : e + a;
        first = 0;
    }
    if (M->terms & 8u) e = first ? M->gauge : e + M->gauge;
    return e;
}

static void marsaglia_f64(or_pcg64* g, double* out) {
    for (;;) {
        double x1 = -1.0 + 2.0 * pcg_f64(g);
        double x2 = -1.0 + 2.0 * pcg_f64(g);
        if (x1 * x1 + x2 * x2 >= 1.0) continue;
        out[0] = 2.0 * x1 * sqrt(1.0 - x1 * x1 - x2 * x2);
        out[1] = 2.0 * x2 * sqrt(1.0 - x1 * x1 - x2 * x2);
        out[2] = 1.0 - 2.0 * (x1 * x1 + x2 * x2);
        return;
    }
}

double or_seq_total_energy(const or_seq* s) {
    const or_model* M = &s->M;
    double total = 0.0;
    if (M->compound) {
        for (uint64_t i = 0; i < s->n; ++i) total += site_energy(s, i);
        return total;
    }
    double si[3], r[3];
    if (M->terms & 1u) {
        for (uint64_t i = 0; i < s->n; ++i) total += e_exchange(s, i);
        return total / 2.0;
    }
    if (M->terms & 2u) {
        ref_vec(s, M->field_dir, r);
        for (uint64_t i = 0; i < s->n; ++i) {
            spin_vec(s, i, si);
            total += sdot(si, r);
        }
        return -fabs(M->field) * total;
    }
    if (M->terms & 4u) {
        ref_vec(s, M->axis, r);
        for (uint64_t i = 0; i < s->n; ++i) {
            spin_vec(s, i, si);
            double d = sdot(si, r);
            total += d * d;
        }
        return total;
    }
    for (uint64_t i = 0; i < s->n; ++i) total += M->gauge;
    return total;
}

double or_seq_magnetization(const or_seq* s) {
    double m[3] = {0.0, 0.0, 0.0}, si[3];
    for (uint64_t i = 0; i < s->n; ++i) {
        spin_vec(s, i, si);
        m[0] += si[0];
        m[1] += si[1];
        m[2] += si[2];
    }
    if (s->ising) return fabs(m[2]);
    double mag = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    return fabs(mag) < DBL_EPSILON ? 0.0 : mag;
}

void or_seq_sweeps(or_seq* s, uint64_t nsweeps, double* energy, double* mag) {
    const double T = s->M.temperature;
    for (uint64_t k = 0; k < nsweeps; ++k) {
        for (uint64_t a = 0; a < s->n; ++a) {
            uint64_t i = pcg_below(&s->rng, s->n);
            double old_e = site_energy(s, i);
            int8_t old_i = 0;
            double old_h[3] = {0.0, 0.0, 0.0};
            if (s->ising) {
                old_i = s->ising[i];
                if (s->M.proposal == 1)
                    s->ising[i] = pcg_f64(&s->rng) < 0.5 ? 1 : -1;
                else
                    s->ising[i] = (int8_t)-old_i;
            } else {
                memcpy(old_h, s->heis + 3 * i, sizeof old_h);
                if (s->M.proposal == 1) {
                    marsaglia_f64(&s->rng, s->heis + 3 * i);
                } else {
                    for (int c = 0; c < 3; ++c) s->heis[3 * i + c] = -old_h[c];
                }
            }
            double delta = site_energy(s, i) - old_e;
            if (delta < 0.0) continue;
            if (pcg_f64(&s->rng) < exp(-delta / T)) continue;
            if (s->ising)
                s->ising[i] = old_i;
            else
                memcpy(s->heis + 3 * i, old_h, sizeof old_h);
        }
        if (energy) energy[k] = or_seq_total_energy(s);
        if (mag) mag[k] = or_seq_magnetization(s);
    }
}

/* ------------------------------------------------------------------------ */
/* Arbitrary exchange matrices: Exchange::new(CsMat<f64>) (energy.rs:170-   */
/* 172, evaluated at 194-213), swept by colour classes.                     */
/*                                                                          */
/* Preparation: the symmetrised pattern (j != i with J_ij or J_ji nonzero)  */
/* is coloured greedily in site order (each site takes the smallest colour  */
/* none of its lower-numbered neighbours has).  Site i's merged row lists   */
/* that pattern by ascending j with a_ij = the J_ij of row i summed in      */
/* entry order and w_ij = a_ij + (J_ji summed likewise) when colour(j) <    */
/* colour(i), else 0; diagonal entries add up to the constant diag.  A      */
/* sweep visits colours 0, 1, ...; the sites of a colour (sorted by id) are */
/* independent.  Per site, in f64 (Ising) or fp32 (Heisenberg), left to     */
/* right: h = sum a_ij s_j and hb = sum w_ij s_j over the merged row; the   */
/* field h gives dE; after the update s_i.hb is the site's share of the     */
/* bond sum sum_{i != j} J_ij s_i.s_j (each pair counted at its            */
/* higher-coloured end, when both spins are final for the sweep).           */
/* ------------------------------------------------------------------------ */
struct or_graph {
    uint64_t n;
    uint32_t ncol;
    uint32_t* colour;   /* per site */
    uint64_t* mptr;     /* merged rows, n + 1 */
    uint64_t* mcol;     /* neighbour site ids, ascending per row */
    double* a;
    double* w;
    double diag;
    uint64_t* order;    /* sites by (colour, id) */
    uint64_t* cstart;   /* ncol + 1 */
};

typedef struct {
    uint64_t row, col, k; /* k: entry index; kind 0 = J_row,col, 1 = transposed */
    uint32_t kind;
    double v;
} or_trip;

static int cmp_trip(const void* pa, const void* pb) {
    const or_trip* a = (const or_trip*)pa;
    const or_trip* b = (const or_trip*)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    if (a->kind != b->kind) return a->kind < b->kind ? -1 : 1;
    return (a->k > b->k) - (a->k < b->k);
}

or_graph* or_graph_create(uint64_t n, const uint64_t* row_ptr, const uint64_t* col,
                          const double* val) {
    or_graph* G = (or_graph*)calloc(1, sizeof(or_graph));
    G->n = n;
    const uint64_t nnz = row_ptr[n];
    or_trip* t = (or_trip*)malloc((2 * nnz + 1) * sizeof(or_trip));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const uint64_t j = col[k];
            if (j == i) {
                G->diag += val[k];
                continue;
            }
            t[m++] = (or_trip){i, j, k, 0, val[k]};
            t[m++] = (or_trip){j, i, k, 1, val[k]};
        }
    qsort(t, (size_t)m, sizeof(or_trip), cmp_trip);
    G->mptr = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
    G->mcol = (uint64_t*)malloc((m + 1) * sizeof(uint64_t));
    G->a = (double*)malloc((m + 1) * sizeof(double));
    G->w = (double*)malloc((m + 1) * sizeof(double));
    double* b = (double*)malloc((m + 1) * sizeof(double));
    uint64_t q = 0;
    for (uint64_t e = 0; e < m; ++e) {
        if (e == 0 || t[e].row != t[e - 1].row || t[e].col != t[e - 1].col) {
            G->mcol[q] = t[e].col;
            G->a[q] = 0.0;
            b[q] = 0.0;
            ++G->mptr[t[e].row + 1];
            ++q;
        }
        if (t[e].kind == 0)
            G->a[q - 1] = G->a[q - 1] + t[e].v;
        else
            b[q - 1] = b[q - 1] + t[e].v;
    }
    for (uint64_t i = 0; i < n; ++i) G->mptr[i + 1] += G->mptr[i];
    free(t);
    /* greedy colouring in site order */
    G->colour = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t maxdeg = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (G->mptr[i + 1] - G->mptr[i] > maxdeg) maxdeg = G->mptr[i + 1] - G->mptr[i];
    uint64_t* stamp = (uint64_t*)calloc(maxdeg + 2, sizeof(uint64_t));
    G->ncol = 0;
    for (uint64_t i = 0; i < n; ++i) {
        for (uint64_t e = G->mptr[i]; e < G->mptr[i + 1]; ++e)
            if (G->mcol[e] < i) stamp[G->colour[G->mcol[e]]] = i + 1;
        uint32_t c = 0;
        while (stamp[c] == i + 1) ++c;
        G->colour[i] = c;
        if (c + 1 > G->ncol) G->ncol = c + 1;
    }
    free(stamp);
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t e = G->mptr[i]; e < G->mptr[i + 1]; ++e)
            G->w[e] = G->colour[G->mcol[e]] < G->colour[i] ? G->a[e] + b[e] : 0.0;
    free(b);
    G->cstart = (uint64_t*)calloc(G->ncol + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) ++G->cstart[G->colour[i] + 1];
    for (uint32_t c = 0; c < G->ncol; ++c) G->cstart[c + 1] += G->cstart[c];
    G->order = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    uint64_t* fill = (uint64_t*)malloc((G->ncol + 1) * sizeof(uint64_t));
    memcpy(fill, G->cstart, (G->ncol + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) G->order[fill[G->colour[i]]++] = i;
    free(fill);
    return G;
}

void or_graph_destroy(or_graph* G) {
    if (!G) return;
    free(G->colour);
    free(G->mptr);
    free(G->mcol);
    free(G->a);
    free(G->w);
    free(G->order);
    free(G->cstart);
    free(G);
}

uint32_t or_graph_colours(const or_graph* G) { return G->ncol; }
void or_graph_colouring(const or_graph* G, uint32_t* colour) {
    memcpy(colour, G->colour, G->n * sizeof(uint32_t));
}
double or_graph_diag(const or_graph* G) { return G->diag; }

/* e^x for x <= 0 in f64 (Cody and Waite's reduction x = n ln2 + r with ln2
 * split so n * hi is exact, then the degree-13 Taylor polynomial of e^r,
 * |r| <= ln2 / 2, coefficients 1/k! rounded to nearest); 0 below -708.
 * Every step is one IEEE operation or an fma, as in the engine. */
double or_exp_f64_nonpos(double x) {
    if (x < -708.0) return 0.0;
    static const double c[14] = {
        0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1,
        0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7,
        0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
        0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
        0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
    const double n = rint(x * 0x1.71547652b82fep+0);
    double r = fma(n, -0x1.62e42fee00000p-1, x);
    r = fma(n, -0x1.a39ef35793c76p-33, r);
    double p = c[13];
    for (int k = 12; k >= 0; --k) p = fma(p, r, c[k]);
    union {
        uint64_t u;
        double d;
    } s;
    s.u = (uint64_t)((int64_t)n + 1023) << 52;
    return p * s.d;
}

void or_graph_ising_init(const or_graph* G, uint64_t seed, int8_t* spins) {
    for (uint64_t i = 0; i < G->n; ++i) {
        uint32_t r[4];
        draw(seed, (uint32_t)i, (uint32_t)(i >> 32), 0, wword(0, 9, 0), r);
        spins[i] = r[0] < 0x80000000u ? 1 : -1;
    }
}

void or_graph_ising_sweeps(const or_graph* G, const or_model* M, uint64_t seed, uint64_t sweep0,
                           uint64_t nsweeps, int8_t* spins, int64_t* sums) {
    const double hm = (M->terms & 2u) ? fabs(M->field) : 0.0;
    const double hz = M->field_dir[2] >= 0.0 ? 1.0 : -1.0;
    const double T = M->temperature < DBL_EPSILON ? DBL_EPSILON : M->temperature;
    const double beta = 1.0 / T;
    for (uint64_t k = 0; k < nsweeps; ++k) {
        const uint64_t t = sweep0 + k;
        int64_t mag = 0, bond = 0;
        for (uint32_t c = 0; c < G->ncol; ++c)
            for (uint64_t p = G->cstart[c]; p < G->cstart[c + 1]; ++p) {
                const uint64_t i = G->order[p];
                uint32_t r[4];
                draw(seed, (uint32_t)i, (uint32_t)(i >> 32), (uint32_t)t, wword(t, 8, 0), r);
                double s = (double)spins[i];
                /* Spin::rand proposes Up when its draw is below 1/2 */
                const int change = M->proposal != 1 || (r[1] < 0x80000000u ? 1.0 : -1.0) != s;
                double h = 0.0, hb = 0.0;
                for (uint64_t e = G->mptr[i]; e < G->mptr[i + 1]; ++e) {
                    const double sj = (double)spins[G->mcol[e]];
                    h = h + G->a[e] * sj;
                    hb = hb + G->w[e] * sj;
                }
                if (change) {
                    const double dE = ((2.0 * s) * h) - ((2.0 * hm) * hz) * s;
                    int accept = dE <= 0.0;
                    if (!accept) accept = (double)r[0] * 0x1p-32 < or_exp_f64_nonpos(-(dE * beta));
                    if (accept) {
                        s = -s;
                        spins[i] = (int8_t)s;
                    }
                }
                mag += (int64_t)s;
                bond += (int64_t)llrint((s * hb) * 134217728.0);
            }
        if (sums) {
            int64_t* o = sums + 5 * k;
            o[0] = bond;
            o[1] = 0;
            o[2] = 0;
            o[3] = mag;
            o[4] = (int64_t)G->n;
        }
    }
}

void or_graph_heis_init(const or_graph* G, uint64_t seed, float* xyz) {
    for (uint64_t i = 0; i < G->n; ++i) {
        float s[3] = {0.0f, 0.0f, 1.0f};
        int done = 0;
        for (uint32_t k = 0; k < 64 && !done; ++k) {
            uint32_t r[4];
            draw(seed, (uint32_t)i, (uint32_t)(i >> 32), k, wword(0, 4, 0), r);
            for (int q = 0; q < 2 && !done; ++q) done = marsaglia_f32(r[2 * q], r[2 * q + 1], s);
        }
        memcpy(xyz + 3 * i, s, sizeof s);
    }
}

void or_graph_heis_sweeps(const or_graph* G, const or_model* M, uint64_t seed, uint64_t sweep0,
                          uint64_t nsweeps, float* xyz, int64_t* sums) {
    const float kk = (M->terms & 4u) ? (float)M->anisotropy : 0.0f;
    const float hm = (M->terms & 2u) ? (float)fabs(M->field) : 0.0f;
    const double T = M->temperature < DBL_EPSILON ? DBL_EPSILON : M->temperature;
    const float beta = (float)(1.0 / T);
    const float axis[3] = {(float)M->axis[0], (float)M->axis[1], (float)M->axis[2]};
    const float fdir[3] = {(float)M->field_dir[0], (float)M->field_dir[1], (float)M->field_dir[2]};
    const float jn = -1.0f; /* the couplings are inside h */
    for (uint64_t k = 0; k < nsweeps; ++k) {
        const uint64_t t = sweep0 + k;
        int64_t acc[5] = {0, 0, 0, 0, 0};
        for (uint32_t c = 0; c < G->ncol; ++c)
            for (uint64_t q = G->cstart[c]; q < G->cstart[c + 1]; ++q) {
                const uint64_t i = G->order[q];
                const uint32_t lo = (uint32_t)i, hi = (uint32_t)(i >> 32);
                float* s = xyz + 3 * i;
                float h[3] = {0.0f, 0.0f, 0.0f}, hb[3] = {0.0f, 0.0f, 0.0f};
                for (int a = 0; a < 3; ++a)
                    for (uint64_t e = G->mptr[i]; e < G->mptr[i + 1]; ++e) {
                        const float sj = xyz[3 * G->mcol[e] + a];
                        const float av = (float)G->a[e], wv = (float)G->w[e];
                        h[a] = h[a] + av * sj;
                        hb[a] = hb[a] + wv * sj;
                    }
                float p[3] = {-s[0], -s[1], -s[2]};
                float u;
                if (M->proposal == 1) {
                    u = random_proposal(seed, lo, hi, t, 0, p);
                } else {
                    uint32_t fw[4];
                    draw(seed, lo, hi, (uint32_t)t, wword(t, 16, 0), fw);
                    u = unit24f(fw[0]);
                }
                const float d[3] = {p[0] - s[0], p[1] - s[1], p[2] - s[2]};
                const float e_ex = jn * dot3f(d, h);
                const float pe = dot3f(p, axis), oe = dot3f(s, axis);
                const float oe2 = oe * oe;
                const float e_an = kk * fmaf(pe, pe, -oe2);
                const float e_ze = hm * dot3f(d, fdir);
                const float e1 = e_ex + e_an;
                const float dE = e1 + e_ze;
                int accept = dE < 0.0f;
                if (!accept) {
                    const float x = -(dE * beta);
                    accept = u < or_exp_nonpos(x);
                }
                if (accept) memcpy(s, p, sizeof p);
                acc[1] += fix27(s[0]);
                acc[2] += fix27(s[1]);
                acc[3] += fix27(s[2]);
                const float se = dot3f(s, axis);
                acc[4] += fix27(se * se);
                acc[0] += (int64_t)llrintf(dot3f(s, hb) * 134217728.0f);
            }
        if (sums) memcpy(sums + 5 * k, acc, sizeof acc);
    }
}

/* The reference's sequential Metropolis over a caller's matrix: the CSR is
 * kept as given (row order, duplicates and diagonal included), as
 * Exchange::new(matrix) evaluates it (energy.rs:194-206). */
or_seq* or_seq_create_csr(const or_model* M, uint64_t n, const uint64_t* row_ptr,
                          const uint64_t* col, const double* val, uint64_t seed) {
    or_seq* s = (or_seq*)calloc(1, sizeof(or_seq));
    s->M = *M;
    s->n = n;
    const uint64_t nnz = row_ptr[n];
    s->indptr = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    s->indices = (uint64_t*)malloc((nnz + 1) * sizeof(uint64_t));
    s->values = (double*)malloc((nnz + 1) * sizeof(double));
    memcpy(s->indptr, row_ptr, (n + 1) * sizeof(uint64_t));
    memcpy(s->indices, col, nnz * sizeof(uint64_t));
    memcpy(s->values, val, nnz * sizeof(double));
    if (M->spin == 2)
        s->heis = (double*)calloc(3 * n, sizeof(double));
    else
        s->ising = (int8_t*)calloc(n, 1);
    or_pcg64_seed(&s->rng, seed);
    return s;
}
