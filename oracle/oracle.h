/*
 * CPU oracle of the MI355X Metropolis engine -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline; the
 * engine (libvegas_hip.so) never links or calls it.
 *
 * Two restatements live here, written independently of the engine's code:
 *
 *  1. The checkerboard semantics the GPU engine defines (Philox4x32-10 keyed
 *     by seed/sweep/colour/row/position, the Ising 32-bit acceptance draw
 *     split into a shared high byte and a per-site low 24 bits, the fp32
 *     Heisenberg update with its per-site Marsaglia draws, the per-sweep
 *     integer sums).  The engine must match it bit for bit.
 *
 *  2. The reference's own sequential algorithm (odarbelaeze/vegas-rs 0.9.0):
 *     N attempts at uniformly random sites per sweep, proposal flip
 *     (integrator.rs:109-138) or fresh random spin (integrator.rs:66-92),
 *     local energy of Exchange over a CSR matrix + Zeeman + Anisotropy +
 *     Gauge in f64 (energy.rs:64-257), acceptance `delta < 0 || u <
 *     exp(-delta/T)` (integrator.rs:127-134), observables by the reference's
 *     total_energy / magnetization conventions (energy.rs:55-59,114-120,
 *     153-160,208-213; state.rs:86-92,150-160,291-296).  Its random stream is
 *     PCG64 (XSL-RR 128/64, the algorithm of rand_pcg::Lcg128Xsl64) but its
 *     seeding and integer/float sampling are not rand 0.9's, so it matches the
 *     reference in distribution, not in trajectory.  Used for 3-sigma parity
 *     and as the CPU baseline.
 *
 * Pins (see tests/): Philox against the Random123 known-answer vectors and
 * rocRAND's host engine; PCG64 against numpy's PCG64; lattice neighbour
 * counts and ground-state energies; the Ising acceptance table against
 * closed forms; observables against exact enumeration of tiny lattices.
 */
#ifndef VEGAS_ORACLE_H
#define VEGAS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_lattice {
    int32_t kind; /* 0 sc, 1 bcc, 2 fcc */
    uint32_t lx, ly, lz;
    uint32_t pbc; /* bit 0 x, bit 1 y, bit 2 z */
} or_lattice;

typedef struct or_model {
    int32_t spin;       /* 0 Ising, 2 Heisenberg */
    int32_t proposal;   /* 0 flip, 1 random */
    uint32_t terms;     /* 1 exchange, 2 zeeman, 4 anisotropy, 8 gauge */
    int32_t compound;   /* Compound total_energy convention */
    double exchange, anisotropy, axis[3], gauge;
    double temperature, field, field_dir[3];
} or_model;

/* Philox4x32-10. */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Lattice: number of sites and the neighbour slots of a site in slot order
 * (sc: -x,+x,-y,+y,-z,+z; bcc: 8 slots; fcc: 12 slots, 4 per other
 * sublattice in ascending order), -1 for an absent neighbour. */
uint64_t or_num_sites(const or_lattice* L);
int or_coordination(const or_lattice* L);
void or_neighbours(const or_lattice* L, uint64_t site, int64_t* slots);
/* colour, row and position of a site in the colour-split planes */
void or_place(const or_lattice* L, uint64_t site, uint32_t* colour, uint32_t* row, uint32_t* pos);

/* Ising acceptance table: thr[b*OR_THR_STRIDE+S], always[b] bit S, S in [0, 2z]. */
#define OR_THR_STRIDE 25
void or_ising_table(const or_lattice* L, const or_model* M, uint32_t* thr, uint32_t* always);

/* Checkerboard semantics.  spins: natural order, int8 +-1 (Ising) or
 * 3 floats per site (Heisenberg).  sums: 5 int64 per sweep
 * {bond, mx, my, mz, axis2} (Ising: axis2 = N, mx = my = 0). */
void or_ising_init(const or_lattice* L, uint64_t seed, int8_t* spins);
void or_ising_sweeps(const or_lattice* L, const or_model* M, uint64_t seed, uint64_t sweep0,
                     uint64_t nsweeps, int8_t* spins, int64_t* sums);
/* The sc Ising checkerboard again, fast and multi-threaded (OpenMP): the
 * CPU baseline's best-CPU line; equal to or_ising_sweeps bit for bit. */
void or_ising_sweeps_sc_fast(const or_lattice* L, const or_model* M, uint64_t seed,
                             uint64_t sweep0, uint64_t nsweeps, int8_t* spins, int64_t* sums,
                             int threads);
void or_heis_init(const or_lattice* L, uint64_t seed, float* xyz);
void or_heis_sweeps(const or_lattice* L, const or_model* M, uint64_t seed, uint64_t sweep0,
                    uint64_t nsweeps, float* xyz, int64_t* sums);
float or_exp_nonpos(float x);

/* Wolff single-cluster steps (integrator.rs:146-230) with the engine's
 * Philox draws; sums as above (bond = each bond once, mag = sum s).  The
 * bond threshold: accept a bond iff its 32-bit draw < thr, or always. */
void or_wolff_threshold(const or_model* M, uint32_t* thr, int* all);
void or_wolff_steps(const or_lattice* L, const or_model* M, uint64_t seed, uint64_t step0,
                    uint64_t nsteps, int8_t* spins, int64_t* sums);

/* Arbitrary exchange matrices (Exchange::new, energy.rs:170-172) swept by
 * the colour classes of a greedy colouring; see oracle.c for the exact
 * semantics.  CSR input: row_ptr[n + 1], col, val (J_ij at row i). */
typedef struct or_graph or_graph;
or_graph* or_graph_create(uint64_t n, const uint64_t* row_ptr, const uint64_t* col,
                          const double* val);
void or_graph_destroy(or_graph* G);
uint32_t or_graph_colours(const or_graph* G);
void or_graph_colouring(const or_graph* G, uint32_t* colour);
double or_graph_diag(const or_graph* G);
double or_exp_f64_nonpos(double x);
/* spins in natural order: Ising int8 +-1, Heisenberg 3 floats per site;
 * sums {bond (2^-27), 0, 0, mag, N} (Ising) or {bond, mx, my, mz, axis2}
 * (Heisenberg, all 2^-27), bond = sum over i != j of J_ij s_i.s_j. */
void or_graph_ising_init(const or_graph* G, uint64_t seed, int8_t* spins);
void or_graph_ising_sweeps(const or_graph* G, const or_model* M, uint64_t seed, uint64_t sweep0,
                           uint64_t nsweeps, int8_t* spins, int64_t* sums);
void or_graph_heis_init(const or_graph* G, uint64_t seed, float* xyz);
void or_graph_heis_sweeps(const or_graph* G, const or_model* M, uint64_t seed, uint64_t sweep0,
                          uint64_t nsweeps, float* xyz, int64_t* sums);

/* PCG64 XSL-RR 128/64: state and increment as (hi, lo) 64-bit halves. */
typedef struct or_pcg64 {
    uint64_t state_hi, state_lo, inc_hi, inc_lo;
} or_pcg64;
void or_pcg64_seed(or_pcg64* g, uint64_t seed);
uint64_t or_pcg64_next(or_pcg64* g);

/* Sequential reference restatement. */
typedef struct or_seq or_seq;
or_seq* or_seq_create(const or_lattice* L, const or_model* M, uint64_t seed);
/* The same over a caller's exchange matrix (CSR kept as given). */
or_seq* or_seq_create_csr(const or_model* M, uint64_t n, const uint64_t* row_ptr,
                          const uint64_t* col, const double* val, uint64_t seed);
void or_seq_destroy(or_seq* s);
void or_seq_set_thermostat(or_seq* s, double temperature, double field, const double dir[3]);
/* spins: Ising int8 +-1 (n) or Heisenberg doubles (3n) */
void or_seq_set_state(or_seq* s, const void* spins);
void or_seq_get_state(const or_seq* s, void* spins);
/* Runs nsweeps sweeps of N attempts; if energy/mag are not NULL, fills the
 * per-sweep reference observables. */
void or_seq_sweeps(or_seq* s, uint64_t nsweeps, double* energy, double* mag);
/* Reference observables of the current state. */
double or_seq_total_energy(const or_seq* s);
double or_seq_magnetization(const or_seq* s);
uint64_t or_seq_nnz(const or_seq* s);

#ifdef __cplusplus
}
#endif

#endif
