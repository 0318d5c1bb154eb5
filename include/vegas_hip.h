/*
 * vegas_hip.h -- C ABI of the MI355X Metropolis engine (libvegas_hip.so).
 *
 * The reference (odarbelaeze/vegas-rs, crate `vegas` 0.9.0) has no FFI: its
 * hot path is the Rust trait call
 *     Integrator<S>::step(&self, rng, &Thermostat<S>, &H, State<S>) -> State<S>
 * (src/integrator.rs:40-49) made once per sweep by Machine::run
 * (src/machine.rs:91-101), followed by every Instrument recomputing
 * H.total_energy and State::magnetization (src/instrument.rs:133-141,254-262).
 * This header is the boundary that replaces that loop: a chain (one lattice,
 * one spin state, one Hamiltonian) lives on the GPU and a call advances it by
 * many sweeps, returning the per-sweep energy and |magnetization| that the
 * reference's instruments would have computed.  Each entry point cites the
 * reference item it replaces.
 *
 * Conventions
 *  - Every function returns an int status (vg_status); 0 is success.  On
 *    failure vg_last_error() returns a thread-local message.
 *  - Handles are opaque and owned by the library; device memory is owned by
 *    the library.  A handle is used by one host thread at a time (the job
 *    server's functions excepted).  Destroy batches and servers before the
 *    context they were created on.
 *  - Host buffers are caller-owned plain arrays.  Site order at the boundary
 *    is the lattice's natural order (vg_lattice, below).
 */
#ifndef VEGAS_HIP_H
#define VEGAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes.  1-6 mirror ProgramError (src/error.rs:30-46), 10      */
/* mirrors VegasError::NotImplementedError (src/error.rs:25-26).        */
/* ------------------------------------------------------------------ */
typedef enum vg_status {
    VG_OK = 0,
    VG_ERR_TEMPERATURE_MAX_LESS_THAN_MIN = 1,
    VG_ERR_NO_STEPS = 2,
    VG_ERR_ZERO_TEMPERATURE = 3,
    VG_ERR_ZERO_COOL_RATE = 4,
    VG_ERR_ZERO_FIELD = 5,
    VG_ERR_ZERO_FIELD_STEP = 6,
    VG_ERR_NOT_IMPLEMENTED = 10,
    VG_ERR_INVALID_ARGUMENT = 20,
    VG_ERR_UNSUPPORTED_LATTICE = 21,
    VG_ERR_DEVICE = 30,
    VG_ERR_OUT_OF_MEMORY = 31,
    VG_ERR_NO_DEVICE = 32,
    VG_ERR_ADMISSION = 40,
    VG_ERR_NOT_FOUND = 41,
    VG_ERR_CANCELLED = 42
} vg_status;

/* ------------------------------------------------------------------ */
/* Lattice.  Replaces the vegas-lattice value built in Input::lattice  */
/* (src/input.rs:296-322: Lattice::sc/bcc(1.0).expand(x,y,z), drop_*)   */
/* and the bench lattice (src/main.rs:25).                              */
/*   sc : site id = x + lx*(y + ly*z); 6 neighbours (+-x,+-y,+-z).      */
/*        lx must be even; a periodic axis must have even length.       */
/*   bcc: site id = 2*(x + lx*(y + ly*z)) + sub; sub 0 = cell corner,   */
/*        sub 1 = body centre; 8 neighbours on the other sublattice.    */
/*   fcc: site id = 4*(x + lx*(y + ly*z)) + sub; sub 0 = corner,        */
/*        subs 1, 2, 3 = the face centres (0,1/2,1/2), (1/2,0,1/2),     */
/*        (1/2,1/2,0); 12 neighbours, 4 on each other sublattice.       */
/* A periodic axis wraps; an open axis (drop_x/y/z) has no bonds across */
/* its ends.  sc and bcc are bipartite (a sweep is two colour passes);  */
/* fcc is not, and a sweep is four passes, one per sublattice.          */
/* ------------------------------------------------------------------ */
#define VG_LATTICE_SC 0
#define VG_LATTICE_BCC 1
#define VG_LATTICE_FCC 2
/* An arbitrary exchange matrix (vg_csr below, vg_batch_create_csr): the  */
/* chain's sites and bonds are the matrix's rows and nonzeros; lx, ly, lz */
/* and pbc are not used.                                                  */
#define VG_LATTICE_GRAPH 3

#define VG_PBC_X 1u
#define VG_PBC_Y 2u
#define VG_PBC_Z 4u

typedef struct vg_lattice {
    int32_t kind;      /* VG_LATTICE_SC, VG_LATTICE_BCC or VG_LATTICE_FCC */
    uint32_t lx, ly, lz; /* unit cells along x, y, z (Lattice::expand) */
    uint32_t pbc;      /* VG_PBC_* bits of the periodic axes */
} vg_lattice;

/* ------------------------------------------------------------------ */
/* Hamiltonian.  Replaces the Hamiltonian value H of energy.rs:         */
/*   Exchange (energy.rs:165-214)      e_i = -J sum_j s_i.s_j           */
/*   Zeeman (energy.rs:125-161)        e_i = +|h| s_i.h_dir             */
/*   UniaxialAnisotropy (84-121)       e_i = +K (s_i.axis)^2            */
/*   Gauge (64-80)                     e_i = g                          */
/*   Compound / hamiltonian! (221-290) e_i = sum of the terms           */
/* The observable energy follows the reference's total_energy exactly:  */
/* a Compound uses the trait default sum_i e_i (energy.rs:55-59) so     */
/* exchange is counted twice; a single bare term uses its own           */
/* total_energy (Exchange: sum/2; Zeeman: -|h| sum s.h_dir;             */
/* Anisotropy: sum (s.axis)^2 without K; Gauge: N g).                   */
/* The Zeeman field itself is a thermostat parameter (thermostat.rs).   */
/* ------------------------------------------------------------------ */
#define VG_TERM_EXCHANGE 1u
#define VG_TERM_ZEEMAN 2u
#define VG_TERM_ANISOTROPY 4u
#define VG_TERM_GAUGE 8u

typedef struct vg_hamiltonian {
    uint32_t terms;             /* VG_TERM_* bits present */
    int32_t compound;           /* 1: Compound of the terms; 0: one bare term */
    double exchange;            /* J, Exchange::from_lattice(exchange, ..) */
    double anisotropy;          /* K, UniaxialAnisotropy::new(_, k) */
    double anisotropy_axis[3];  /* reference spin of the anisotropy */
    double gauge;               /* Gauge::new(value) */
} vg_hamiltonian;

/* ------------------------------------------------------------------ */
/* Chain = lattice + spin kind + proposal + Hamiltonian + seed.          */
/* ------------------------------------------------------------------ */
#define VG_SPIN_ISING 0          /* IsingSpin (state.rs:61-130), int8 arena */
#define VG_SPIN_ISING_BITS 1     /* IsingSpin, 1 bit per site arena (every lattice) */
#define VG_SPIN_HEISENBERG 2     /* HeisenbergSpin (state.rs:134-192), fp32 xyz */

#define VG_PROPOSAL_FLIP 0       /* MetropolisFlipIntegrator (integrator.rs:100-138) */
#define VG_PROPOSAL_RANDOM 1     /* MetropolisIntegrator (integrator.rs:57-92) */
/* WolffIntegrator (integrator.rs:146-230), int8 Ising chains only: one   */
/* "sweep" of vg_batch_run is one Wolff step, i.e. one cluster grown from */
/* a random site with bond probability 1 - exp(-2J/T) over equal spins    */
/* and flipped (the move ignores the field, as the reference's does).     */
/* Chains of fewer than 2^31 sites (VG_ERR_INVALID_ARGUMENT otherwise).   */
#define VG_PROPOSAL_WOLFF 2

typedef struct vg_chain_desc {
    vg_lattice lattice;
    int32_t spin;               /* VG_SPIN_* */
    int32_t proposal;           /* VG_PROPOSAL_* */
    vg_hamiltonian hamiltonian;
    uint64_t seed;              /* Philox key of this chain */
} vg_chain_desc;

typedef struct vg_ctx vg_ctx;
typedef struct vg_batch vg_batch;

/* Exchange::new(CsMat<f64>) (src/energy.rs:170-172): the exchange of a     */
/* VG_LATTICE_GRAPH chain as a sparse matrix J in CSR form, row i holding  */
/* the J_ij of site i, evaluated as the reference does (energy.rs:194-213):*/
/*   e_i = -sum_j J_ij s_i.s_j over row i; Exchange's total_energy halves   */
/*   sum_i e_i; a Compound adds the e_i of every term.                     */
/* The couplings live in the matrix (vg_hamiltonian.exchange is not used; */
/* the VG_TERM_EXCHANGE bit must be set).  Duplicate entries of a row are */
/* summed; a diagonal entry J_ii adds the constant -J_ii to e_i.  The     */
/* sweep visits the colour classes of a greedy colouring of the matrix's  */
/* symmetrised pattern in turn, all sites of a class at once (sites of one */
/* class share no nonzero); vg_batch_colours reports the class count.     */
typedef struct vg_csr {
    uint64_t n_sites;          /* rows = columns = sites, 1 .. 2^31 - 1 */
    const uint64_t* row_ptr;   /* n_sites + 1 offsets, row_ptr[0] = 0 */
    const uint64_t* col;       /* row_ptr[n_sites] column indices (< n_sites) */
    const double* value;       /* J_ij, finite */
} vg_csr;

/* Per-sweep integer sums the kernels reduce (fixed point for Heisenberg,
 * scale vg_sums_scale()).  bond = sum over bonds of s_i.s_j counted once;
 * mag = sum_i s_i; axis2 = sum_i (s_i.axis)^2. */
typedef struct vg_sweep_sums {
    int64_t bond;
    int64_t mag[3];
    int64_t axis2;
} vg_sweep_sums;

/* ---------------------------- context ------------------------------ */
int vg_device_count(int* count);
int vg_ctx_create(int device, vg_ctx** out);
int vg_ctx_destroy(vg_ctx* ctx);
const char* vg_last_error(void);
const char* vg_version(void);
/* 1/scale of the fixed-point sums in vg_sweep_sums for a spin kind. */
double vg_sums_scale(int32_t spin);
/* The kernels' Philox4x32-10 evaluated on the host (same source), for
 * inspection: out = Philox(counter, key). */
void vg_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);

/* ------------------------------------------------------------------ */
/* Batch of independent chains packed in one device arena.  One chain is */
/* a batch of size 1.  Replaces Machine<H,I,S> (machine.rs:44-125) for   */
/* each chain: the state stays on the device between calls.              */
/* ------------------------------------------------------------------ */
int vg_batch_create(vg_ctx* ctx, uint32_t n_chains, const vg_chain_desc* descs,
                    vg_batch** out);
/* vg_batch_create for batches holding VG_LATTICE_GRAPH chains (Exchange::new
 * with an arbitrary matrix, energy.rs:170-172): exchange[i] is the matrix of
 * chain i when its lattice kind is VG_LATTICE_GRAPH and is ignored otherwise
 * (lattice chains may share the batch).  Graph chains take int8 Ising
 * (flip or random proposals) or Heisenberg spins; the matrix is copied. */
int vg_batch_create_csr(vg_ctx* ctx, uint32_t n_chains, const vg_chain_desc* descs,
                        const vg_csr* exchange, vg_batch** out);
/* Colour classes (passes per sweep) of a chain: 2 for sc and bcc, 4 for fcc,
 * the greedy colouring's count for a graph chain. */
int vg_batch_colours(const vg_batch* batch, uint32_t chain, uint32_t* colours);
/* Scale of each of the five vg_sweep_sums quantities of a chain (value =
 * integer * scale).  Lattice chains: vg_sums_scale(spin) for all five.
 * Graph chains: bond = sum over the off-diagonal nonzeros of J_ij s_i.s_j
 * (the couplings included) at 2^-27; mag and axis2 as for lattice chains of
 * the same spin kind. */
int vg_batch_sums_scales(const vg_batch* batch, uint32_t chain, double scales[5]);
int vg_batch_destroy(vg_batch* batch);
int vg_batch_size(const vg_batch* batch, uint32_t* n_chains);
int vg_batch_num_sites(const vg_batch* batch, uint32_t chain, uint64_t* n_sites);
/* Machine::set_thermostat / Thermostat::new (thermostat.rs:29-39, 50-68):
 * temperature below DBL_EPSILON is clamped to DBL_EPSILON; the field
 * magnitude is used as |magnitude| (Field::magnitude, state.rs:219-221);
 * orientation is the field's spin (Ising: sign of orientation[2]). */
int vg_batch_set_thermostat(vg_batch* batch, uint32_t chain, double temperature,
                            double field_magnitude, const double field_orientation[3]);
/* State::rand_with_size (state.rs:260-262): i.i.d. spins from the chain's
 * seed (Ising: Up with probability 1/2; Heisenberg: uniform on S^2 by
 * Marsaglia's method, util.rs:21-34).  Resets the chain's sweep counter. */
int vg_batch_randomize(vg_batch* batch);
/* State upload / download in natural site order.  Ising: int8 +1 (Up) or
 * -1 (Down).  Heisenberg: 3 doubles per site, stored as fp32 on device. */
int vg_batch_set_ising(vg_batch* batch, uint32_t chain, const int8_t* spins);
int vg_batch_get_ising(vg_batch* batch, uint32_t chain, int8_t* spins);
int vg_batch_set_heisenberg(vg_batch* batch, uint32_t chain, const double* xyz);
int vg_batch_get_heisenberg(vg_batch* batch, uint32_t chain, double* xyz);
int vg_batch_sweep_count(const vg_batch* batch, uint32_t chain, uint64_t* sweeps);
int vg_batch_set_sweep_count(vg_batch* batch, uint32_t chain, uint64_t sweeps);
/* Run `sweeps` sweeps of every chain (Machine::run, machine.rs:91-101, with
 * Integrator::step = one checkerboard sweep of N single-site attempts).
 * energy / magnetization, if not NULL, receive n_chains*sweeps values,
 * chain-major: the total energy (Hamiltonian::total_energy convention, see
 * vg_hamiltonian) and |sum_i s_i| (State::magnetization().magnitude(),
 * state.rs:291-296) after each sweep -- what StatSensor/ObservableSensor
 * record (instrument.rs:133-141, 254-262). */
int vg_batch_run(vg_batch* batch, uint64_t sweeps, double* energy, double* magnetization);
/* Same sweeps, returning the raw per-sweep integer sums (n_chains*sweeps). */
int vg_batch_run_sums(vg_batch* batch, uint64_t sweeps, vg_sweep_sums* sums);
/* Execution options: 1 = capture sweep blocks into a hipGraph and replay. */
int vg_batch_set_graph(vg_batch* batch, int enabled);
/* With graphs enabled: captures, instantiates and uploads the graphs that a
 * run of `sweeps` sweeps replays (a block of the batch's full length and
 * one of the remainder), without running any sweep, so that the run itself
 * only launches.  No counterpart in the reference (its sweeps are plain
 * loops); for latency-sensitive callers and timing.  A layout change (tile,
 * autotune) drops the captured graphs. */
int vg_batch_prepare(vg_batch* batch, uint64_t sweeps);
/* Tile of the Ising sc, bcc and fcc passes (int8 and bit-packed): z
 * planes marched by one work item (1, 2, 4 or 8; default 4), set for all
 * six groups; get_tile reports the first of them that has chains.  Results
 * do not depend on it. */
int vg_batch_set_tile(vg_batch* batch, uint32_t tile);
int vg_batch_get_tile(const vg_batch* batch, uint32_t* tile);
/* Times `trial_sweeps` sweeps with every tile (replays of graphs captured
 * beforehand, five rounds over the tiles, the median of each) and keeps the
 * fastest, for each of the six groups.  The chains' spins and sweep counters are
 * restored after every trial, so tuning does not change any trajectory. */
int vg_batch_autotune(vg_batch* batch, uint32_t trial_sweeps, uint32_t* chosen_tile);
/* Records (0..128, default 128) the per-wave tie queue of the Ising sc, bcc
 * and fcc passes may hold before a wave settles further tied rows in the
 * lane that found them (each pass holds at most its own queue: 128 records
 * in the sc pass, 64 in the bcc and fcc passes).  No counterpart in the reference; a test hook, so
 * that the queue-overflow path can be checked against the oracle: results
 * do not depend on it.  Drops captured graphs. */
int vg_batch_set_tie_capacity(vg_batch* batch, uint32_t records);
/* Last run's device time in milliseconds (hipEvent pair around the sweeps). */
int vg_batch_last_run_ms(const vg_batch* batch, double* ms);
/* Per-launch timing of the colour-pass kernels: when enabled (plain launches
 * only), every pass launch is bracketed by a hipEvent pair on the batch
 * stream; vg_batch_pass_stats returns the accumulated kernel time and launch
 * count of the pass kernel used by `spin` chains.  Enabling resets them. */
int vg_batch_set_timing(vg_batch* batch, int enabled);
int vg_batch_pass_stats(const vg_batch* batch, int32_t spin, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------ */
/* Allocation-free sweeps.  No counterpart in the reference: for callers */
/* that cannot allocate device memory (code under HIP graph capture,     */
/* tight serving loops, harnesses that forbid hipMalloc).  The sweeps of  */
/* vg_batch_run_sums for one Metropolis Ising chain (int8 or bit-packed;  */
/* sc, bcc or fcc; tile 4) on caller-owned device buffers, enqueued on    */
/* `stream` (a hipStream_t; NULL = the null stream) of the calling        */
/* thread's current device; the call returns once the work is enqueued.  */
/*   spins: the chain's colour planes, plane c (c < 2, or < 4 on fcc) at  */
/*     byte offset c * round_up(ly * lz * pitch, 256); row r = y + ly * z */
/*     at r * pitch; position i of a row holds the colour-c site at       */
/*     x = 2i + ((c + y + z) & 1) on sc, the sublattice-c site of cell    */
/*     x = i on bcc and fcc; w = lx / 2 (sc) or lx positions.  int8: one  */
/*     byte per position, 1 = Up, 0 = Down, pitch = round_up(w, 64).      */
/*     Bit-packed: bit (i & 7) of byte i >> 3, pitch = round_up(2 *       */
/*     ceil(w / 16), 16).  Padding bytes zero.  Updated in place.          */
/*   sums: sweeps * 5 int64, per sweep {bond, 0, 0, mag, 0} as in          */
/*     vg_sweep_sums (the batch path reports N as the fifth).             */
/*   sweep0: the chain's sweep counter before the first sweep (Philox).   */
/* The thermostat follows vg_batch_set_thermostat.  Descriptor, block map */
/* and pass partials live in a 16 MiB device array that each GPU holds    */
/* until the process ends.  A sweep's partials take colours x workgroups  */
/* x 32 bytes (N / 512 bytes on sc, cells / 128 on fcc); a chain whose    */
/* partials for one sweep do not fit is refused (VG_ERR_INVALID_ARGUMENT):*/
/* every sc and bcc chain a batch accepts fits, fcc up to about 2 x 10^9  */
/* cells.                                                                 */
/* Calls on one GPU share that array: calls that overlap on one GPU race */
/* (serialise them, e.g. on one stream).  Nothing needs clearing between */
/* calls: a call writes all it reads.                                     */
int vg_run_sums_noalloc(const vg_chain_desc* desc, double temperature, double field_magnitude,
                        const double field_orientation[3], uint64_t sweep0, uint64_t sweeps,
                        void* spins, int64_t* sums, void* stream);

/* ------------------------------------------------------------------ */
/* Job server: continuous batching of simulation jobs on one GPU.       */
/* A job is a chain plus a program -- the stages a Program::run drives   */
/* (Relax, CoolDown, HysteresisLoop, program.rs:97-336) -- and its      */
/* result is what StatSensor writes per measured stage                  */
/* (instrument.rs:93-142).  No counterpart in the reference (which runs */
/* one Machine at a time); it serves many Machines at once.             */
/*                                                                      */
/* Jobs are admitted in arrival order (FIFO) when a slot is free and    */
/* their spins fit in the server's device pool (admission by HBM        */
/* footprint); admitted jobs share every sweep block with the jobs      */
/* already running.  Each job's spins, sweep counter (its Philox        */
/* counter) and accumulators stay resident until it finishes, and its   */
/* results are bit-identical whether it runs alone or in any mix.       */
/* The server runs its own worker thread and HIP stream; the functions  */
/* below may be called from any thread.                                 */
/* ------------------------------------------------------------------ */
typedef struct vg_server vg_server;

typedef struct vg_server_config {
    uint64_t memory_bytes;   /* device pool holding the admitted jobs' spins */
    uint32_t max_jobs;       /* jobs on the device at once (slots) */
    uint32_t quantum_sweeps; /* sweeps per scheduling quantum, 1..256 */
    uint32_t tile;           /* Ising sc / bcc / fcc tile: 1, 2, 4 or 8 */
    int32_t use_graph;       /* 1: replay quanta from hipGraphs while the job set is unchanged */
    uint64_t sweep_budget;   /* work per sweep of a quantum, in Ising-site units (a Heisenberg
                                site counts 15): jobs with the least remaining work are taken
                                first up to this sum, the rest wait for a later quantum;
                                0 = every running job in every quantum */
} vg_server_config;

/* One stage: set the thermostat, run `relax` sweeps, then `measure`
 * sweeps whose E and |M| are accumulated.  The thermostat is
 * Thermostat::new(temperature, Field::new(orientation, field)) with the
 * reference's semantics (T clamped to DBL_EPSILON, |field| applied). */
typedef struct vg_stage {
    double temperature;
    double field;
    double field_dir[3];
    uint64_t relax;
    uint64_t measure;
} vg_stage;

typedef struct vg_job_desc {
    vg_chain_desc chain;     /* initial spins: State::rand_with_size from chain.seed */
    uint32_t n_stages;
    const vg_stage* stages;
} vg_job_desc;

/* StatSensor's line for a stage with measure > 0 (instrument.rs:113-123):
 * T, |h|, <E>, Var(E)/(N T^2), <|M|>, Var(|M|)/(N T), Binder(|M|). */
typedef struct vg_stage_result {
    double temperature;
    double field;
    double energy;
    double specific_heat;
    double magnetization;
    double susceptibility;
    double binder;
    uint64_t measured;       /* sweeps accumulated (0: relax-only stage) */
} vg_stage_result;

#define VG_JOB_QUEUED 0
#define VG_JOB_RUNNING 1
#define VG_JOB_DONE 2
#define VG_JOB_FAILED 3
#define VG_JOB_CANCELLED 4

typedef struct vg_job_status {
    int32_t state;           /* VG_JOB_* */
    uint32_t stage;          /* current stage index */
    uint64_t sweeps_done;
    uint64_t sweeps_total;
    double submit_s;         /* seconds since the server started; -1 = not yet */
    double start_s;          /* admitted */
    double end_s;            /* finished */
    double last_energy;      /* E and |M| after the job's latest sweep */
    double last_magnetization;
} vg_job_status;

typedef struct vg_server_stats {
    uint64_t submitted, completed, running, queued;
    uint64_t attempts;       /* spin-flip attempts executed */
    uint64_t quanta;         /* sweep blocks launched */
    double device_ms;        /* device time of the sweep blocks */
    uint64_t pool_bytes, pool_used, pool_peak;
    uint32_t peak_jobs;      /* most jobs on the device at once */
    uint32_t pad;
} vg_server_stats;

int vg_server_create(vg_ctx* ctx, const vg_server_config* config, vg_server** out);
/* Cancels queued and running jobs, stops the worker, frees the pool. */
int vg_server_destroy(vg_server* server);
/* Queues a job.  VG_ERR_ADMISSION if its spins could never fit the pool;
 * program errors (no stages, zero steps) as the reference's ProgramError. */
int vg_server_submit(vg_server* server, const vg_job_desc* job, uint64_t* job_id);
int vg_server_status(vg_server* server, uint64_t job_id, vg_job_status* status);
/* Blocks until the job is done, failed or cancelled, or timeout_s passes
 * (timeout_s < 0: no limit); fills status either way. */
int vg_server_wait(vg_server* server, uint64_t job_id, double timeout_s, vg_job_status* status);
/* Per-stage results of a finished job (n = its n_stages). */
int vg_server_results(vg_server* server, uint64_t job_id, vg_stage_result* results, uint32_t n);
int vg_server_cancel(vg_server* server, uint64_t job_id);
/* Forgets a finished job (its status and results). */
int vg_server_release(vg_server* server, uint64_t job_id);
int vg_server_get_stats(vg_server* server, vg_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* VEGAS_HIP_H */
