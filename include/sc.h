/*
 * sc.h -- C-ABI of libsc, the MI355X (gfx950) molecular-dynamics hot path of
 * SimulationCode.jl.
 *
 * Plain pointers and sizes only.  Host arrays use the reference's host layout:
 * Julia's Vector{SVector{dims,Float64}} is a contiguous AoS block of
 * N x dims doubles ([x1 y1 z1 x2 y2 z2 ...], src/DataStructs.jl:183-201);
 * diameters are N doubles.  Host arrays are borrowed for the duration of a
 * call only (copied synchronously); the library keeps no host pointers.
 * Device state is owned by the context and kept resident between calls.
 *
 * Every entry point returns an int status (SC_OK == 0, negative on error);
 * sc_last_error() gives the text.  A context is not thread-safe; independent
 * contexts are.  Each context owns one HIP stream (default priority).
 *
 * Reference entry points each function replaces are cited as file:line
 * relative to the SimulationCode.jl checkout.
 */
#ifndef SC_H
#define SC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SC_POT_LJ = 1, SC_POT_BERTHIER = 2 };
enum { SC_SYS_NEWTONIAN = 0, SC_SYS_LANGEVIN = 1, SC_SYS_BROWNIAN = 2 };
/* SC_PREC_FP64: all arithmetic in FP64 (the drop-in).
 * SC_PREC_MIXED: pair distance and force arithmetic in FP32; positions,
 * velocities, minimum-image wrapping, integrator updates and energy
 * reductions stay FP64. */
enum { SC_PREC_FP64 = 0, SC_PREC_MIXED = 1 };

enum {
    SC_OK = 0,
    SC_ERR_ARG = -1,       /* bad argument / unsupported combination        */
    SC_ERR_OOM = -2,       /* device allocation failed                      */
    SC_ERR_OVERFLOW = -3,  /* a full-list row exceeded maxneighbours        */
    SC_ERR_NONFINITE = -4, /* a position became NaN/Inf                     */
    SC_ERR_HIP = -5,       /* HIP runtime error                             */
    SC_ERR_STATE = -6      /* call out of order (e.g. no neighbour lists)   */
};

/* ParameterStruct + System + InteractionPotential + the scalar arguments of
 * initialize_neighbor_struct (src/DataStructs.jl:10-19, 45-87, 138-153,
 * 349-354).  Field meanings follow the reference. */
typedef struct sc_params {
    int32_t dims;                 /* 2 or 3 */
    int32_t potential;            /* SC_POT_* */
    int32_t system;               /* SC_SYS_* */
    int32_t precision;            /* SC_PREC_* */
    int32_t rescale_temperature;  /* Newtonian.rescale_temperature */
    int32_t cell_divisor;         /* cells of side >= R_v / cell_divisor, stencil +-cell_divisor; 0 => 2 */
    int32_t build_mode;           /* 0: automatic (the per-slot sweep testing candidates on fixed-point
                                     records in FP32, the FP64 predicates deciding within its error
                                     margins; the FP64 sweep where the margins do not hold);
                                     1: LDS-staged cell runs (cell_divisor 1); 2: per-slot FP64 stencil
                                     sweep; 3: LDS windows with an FP32 prefilter (cell_divisor 1-3);
                                     4: LDS-staged stencil rows, one wavefront per 64 slots; every mode
                                     writes the same lists */
    int32_t reserved;             /* must be 0 */
    int64_t N;
    int64_t maxneighbours;        /* initialize_neighbor_struct kwarg (200) */
    double box_size;              /* ParameterStruct.box_size */
    double force_cutoff;          /* initialize_neighbor_struct force_cutoff */
    double force_cutoff2;         /* ParameterStruct.force_cutoff2 */
    double skin_distance;         /* initialize_neighbor_struct skin_distance */
    double lj_eps, lj_sigma, lj_Uc;          /* LJ (DataStructs.jl:10-19) */
    double b_c0, b_c2, b_c4, b_zeta;          /* Berthier (DataStructs.jl:45-60) */
    double kBT, gamma, dt, m;                 /* System (DataStructs.jl:64-87) */
    uint64_t seed;                /* on-device Philox4x32-10 noise seed */
} sc_params;

typedef struct sc_ctx sc_ctx;

/* initialize_neighbor_struct(skin, box_size, force_cutoff, N, D, dims;
 * maxneighbours) -- src/DataStructs.jl:349-378.  Allocates the device
 * state for one system; R_v = skin + force_cutoff * max(D).  Every
 * diameter must be finite and positive (SC_ERR_ARG otherwise, before any
 * device work; the same holds for the D of sc_set_state and sc_batch_add). */
int sc_create(sc_ctx **out, const sc_params *p, const double *D);
void sc_destroy(sc_ctx *ctx);
const char *sc_last_error(const sc_ctx *ctx);

/* Upload/download ArrayStruct fields (src/DataStructs.jl:183-201) and the
 * NeighborListStruct displacement arrays (:252-253).  Any pointer may be
 * NULL (field untouched).  Setting r does not rebuild the lists, exactly as
 * assigning arrays.r_array in the reference does not. */
int sc_set_state(sc_ctx *ctx, const double *r, const double *v, const double *F,
                 const double *D, const double *disp);
int sc_get_state(sc_ctx *ctx, double *r, double *v, double *F, double *disp,
                 double *disp2);

/* update_neighbor_lists! -- src/NeighborLists.jl:24-53 (counts a rebuild,
 * zeroes the displacements, rebuilds the Verlet lists on the device).  It
 * also clears an error a previous call left on the device (see
 * sc_do_time_steps): the rebuild checks the lists again. */
int sc_update_neighbor_lists(sc_ctx *ctx);

/* Reference-layout lists for parity and for the CPU Monte-Carlo path
 * (NeighborListStruct fields, src/DataStructs.jl:248-251): counts of length
 * N, lists column-major maxneigh x N, 1-based indices.  As in the reference,
 * both lists are the ones the last rebuild stored (src/NeighborLists.jl:82-104),
 * evaluated at the positions of that rebuild, whatever the particles did
 * since: full rows hold every j with d2 <= R_v^2; half rows hold each pair
 * with d2 < (sqrt(force_cutoff2)*sigma_ij + skin)^2 once (in the row of the
 * smaller index).  Entry order within a row is unspecified, as in the
 * reference.  Any output may be NULL. */
int sc_get_neighbor_lists(sc_ctx *ctx, int64_t maxneigh, int64_t *nfull, int64_t *full,
                          int64_t *nhalf, int64_t *half);

/* The lists of sc_get_neighbor_lists (same layout and entries), for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving or Monte-Carlo loops -- with no counterpart in the reference.  It
 * calls no hipMalloc/hipFree: the rows are assembled in a 64 MiB scratch
 * array of static device memory (each GPU the library runs on holds it until
 * the process ends) and copied out.  Entries past each row's count are set
 * to 0.  Each requested list needs N * (maxneigh + 1) 8-byte words of the
 * scratch; beyond 8388576 words in total the call returns SC_ERR_ARG.  Calls
 * on contexts of one GPU share the scratch: calls that overlap race. */
int sc_get_neighbor_lists_noalloc(sc_ctx *ctx, int64_t maxneigh, int64_t *nfull, int64_t *full,
                                  int64_t *nhalf, int64_t *half);

/* calculate_new_forces! -- src/MD.jl:262-316, including
 * add_new_random_force! (:30-82).  noise: N x dims AoS standard normals
 * (what randn! would have produced), or NULL for on-device Philox noise. */
int sc_calculate_new_forces(sc_ctx *ctx, const double *noise);

/* nsteps calls of do_time_step -- src/MD.jl:499-554 (Newtonian/Langevin
 * velocity Verlet with COM removal, optional rescale, skin-triggered
 * rebuild) and the intended Brownian step of src/MD.jl:429-471.  noise:
 * nsteps consecutive N x dims AoS blocks, or NULL for on-device Philox.
 * The state is reference-consistent when the call returns.  A device error
 * (overflow, non-finite position) stops the call with its status and stays
 * recorded: this call, sc_calculate_new_forces and the minimisers then refuse
 * with the same status until sc_update_neighbor_lists has run. */
int sc_do_time_steps(sc_ctx *ctx, int64_t nsteps, const double *noise);

/* calculate_full_energy -- src/AuxiliaryFunctions.jl:22-34 (+ find_energy
 * :546-584): sum over full-list pairs, halved; compensated reduction. */
int sc_full_energy(sc_ctx *ctx, double *out);
/* compute_potential_energy! -- src/AuxiliaryFunctions.jl:186-196: E[i] =
 * find_energy(i) / 2 for every particle (original order, N doubles); the
 * potential_energy_array that save_data dumps (src/Dump.jl:82-85). */
int sc_potential_energies(sc_ctx *ctx, double *E);
/* calculate_full_energy_no_neigh -- src/AuxiliaryFunctions.jl:86-110: the
 * O(N^2) all-pairs energy, tiled on the device. */
int sc_energy_no_neigh(sc_ctx *ctx, double *out);
/* find_kinetic_energy -- src/AuxiliaryFunctions.jl:57-65. */
int sc_kinetic_energy(sc_ctx *ctx, double *out);
/* rescale_velocities! -- src/MD.jl:391-402. */
int sc_rescale_velocities(sc_ctx *ctx);

/* arrays.r_old_array .= arrays.r_array -- src/MD.jl:682 (start of
 * perform_molecular_dynamics!): the reference configuration of
 * sc_f2_self, kept per original particle index. */
int sc_set_reference_positions(sc_ctx *ctx);
/* calculate_F2_self! -- src/AuxiliaryFunctions.jl:968-987: mean over
 * particles of cos(kx x_old) cos(kx x) + sin(kx x_old) sin(kx x) on the x
 * components (the reference uses kx = 7.2); SC_ERR_STATE before
 * sc_set_reference_positions. */
int sc_f2_self(sc_ctx *ctx, double kx, double *F2s);
/* calculate_ql (averaged = 0) -- src/AuxiliaryFunctions.jl:687-711 with
 * calculate_qlm :758-801 -- and calculate_ql_averaged (averaged = 1,
 * :832-902): the Steinhardt order parameter of order l (1..8), averaged
 * over particles, from neighbours within q_cutoff (at most the Verlet
 * radius: the reference takes them from the full list), with the
 * reference's angles (theta = atan(sqrt(x^2+y^2)/z), phi = atan(y/x)).
 * 3D only (the reference asserts dims == 3): SC_ERR_ARG otherwise. */
int sc_steinhardt(sc_ctx *ctx, int32_t l, double q_cutoff, int32_t averaged, double *ql);

/* OutputStruct.N_neighbor_list_rebuilds (src/DataStructs.jl:299): read and
 * optionally overwrite (pass set >= 0). */
int sc_rebuild_count(sc_ctx *ctx, int64_t *out, int64_t set);

/* find_inherent_structure's capped steepest-descent step
 * (src/InherentStructure.jl:26-51): `iters` iterations of
 * r += min(1, max_step/max|F*dt|) * F*dt with the skin-triggered rebuild.
 * Returns the final max |F| in *maxforce (may be NULL). */
int sc_minimize(sc_ctx *ctx, int64_t iters, double dt, double max_step, double *maxforce);

/* find_inherent_structure's loop itself (src/InherentStructure.jl:19-56): the
 * iterations of sc_minimize until max(max|F|, max|F*dt|) <= tolerance, the
 * check made on the device after each iteration's move and rebuild, as the
 * reference makes it; at most max_iters iterations.  *iterations gets the
 * count done (the reference's `iteration`), *maxforce the max |F| of the last
 * one and *converged 1 if the tolerance was met (0: the reference raises
 * "Not converged").  F afterwards holds the last iteration's forces.  Any of
 * the outputs may be NULL. */
int sc_minimize_until(sc_ctx *ctx, int64_t max_iters, double dt, double max_step, double tolerance,
                      int64_t *iterations, double *maxforce, int32_t *converged);

/* The on-device noise generator's counter-based block function
 * (Philox4x32-10, Salmon et al., SC'11), exported so its known-answer
 * vectors can be checked on the host.  No reference counterpart: the
 * reference draws noise from Julia's global RNG (src/MD.jl:4-6). */
void sc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Device-side wall time of the last sc_do_time_steps call (HIP events on the
 * context stream), in milliseconds. */
int sc_last_elapsed_ms(sc_ctx *ctx, double *ms);
/* Measurement hook, no reference counterpart: the device's streaming
 * ceiling, a plain copy of `bytes` (read + write counted) timed over `reps`
 * launches; allocates and frees its own buffers. */
int sc_copy_bandwidth(int64_t bytes, int32_t reps, double *gbps);

/* Measurement hooks (no reference counterpart; the reference only wraps its
 * loop in @time, src/MD.jl:704).  With timing on, sc_do_time_steps brackets
 * every pair-force launch and every drift launch with HIP events on the
 * context stream; sc_kernel_times returns the accumulated milliseconds and
 * launch counts since timing was switched on. */
int sc_set_kernel_timing(sc_ctx *ctx, int on);
int sc_kernel_times(sc_ctx *ctx, double *force_ms, int64_t *force_launches, double *drift_ms,
                    int64_t *drift_launches);

/* Short-list statistics of the current lists: total entries (each pair
 * counted from both sides) and the longest row. */
int sc_neighbor_stats(sc_ctx *ctx, int64_t *total_entries, int64_t *max_row);

/* ------------------------------------------------------------------------
 * Replica batches (no reference counterpart: the reference runs one system
 * per Julia process).  Independent systems of any N, dims, potential,
 * dynamics and precision are admitted into a batch and advanced together,
 * each kernel of a step launched once for all of them; the step may be
 * captured into a hipGraph.  Replicas draw on-device Philox noise keyed by
 * their own sc_params.seed, so a replica's trajectory is bitwise identical
 * alone and in any mix.  Each replica follows the reference step of its own
 * dynamics (src/MD.jl:429-554). */
typedef struct sc_batch sc_batch;

int sc_batch_create(sc_batch **out, int32_t max_replicas);
void sc_batch_destroy(sc_batch *b);
const char *sc_batch_last_error(const sc_batch *b);
/* Device bytes a replica with these parameters and diameters would occupy
 * (for admission by HBM budget); -1 on bad arguments. */
int64_t sc_replica_device_bytes(const sc_params *p, const double *D);
/* Admit a replica: allocate its state, upload r (and v, may be NULL), build
 * its lists.  *id identifies it until sc_batch_remove. */
int sc_batch_add(sc_batch *b, const sc_params *p, const double *D, const double *r, const double *v,
                 int32_t *id);
int sc_batch_remove(sc_batch *b, int32_t id);
/* Capture the step into a hipGraph (re-captured when the membership
 * changes) and replay it, instead of launching kernel by kernel. */
int sc_batch_set_graph(sc_batch *b, int32_t on);
/* Advance every admitted replica by nsteps steps.  By default the steps are
 * enqueued ahead in chunks: a replica whose skin trigger fires halts for the
 * rest of the chunk while the others go on, and the replicas that halted in a
 * chunk are rebuilt together after it (one rebuild chain per chunk instead of
 * one per step); every replica's trajectory is bitwise the one of eager
 * stepping.  Batches with rescale_temperature replicas or a captured graph
 * step eagerly. */
int sc_batch_step(sc_batch *b, int64_t nsteps);
/* Steps launched ahead per chunk (default 8), or 0 for one launch sequence
 * per step; the results are the same. */
int sc_batch_set_ahead(sc_batch *b, int32_t chunk);
/* sc_batch_step in two halves: start the steps and return at once (several
 * batches then run concurrently on their own streams; the chunked loop runs
 * on a host thread of the library), and wait for them (device time, step
 * counters, error check).  Any other call on the batch waits first, except
 * sc_batch_last_error, whose message the steps may still set: read it after
 * sc_batch_wait.  Calls on one batch from several host threads at once are
 * not supported (different batches are independent). */
int sc_batch_step_async(sc_batch *b, int64_t nsteps);
int sc_batch_wait(sc_batch *b);
/* Capped steepest descent (sc_minimize) of every admitted replica. */
int sc_batch_minimize(sc_batch *b, int64_t iters, double dt, double max_step);
/* sc_minimize_until for every admitted replica at once (find_inherent_structure,
 * src/InherentStructure.jl:19-56, per replica): each replica stops on its own
 * iteration count, exactly as it would alone.  iterations[id] and
 * converged[id] (arrays of max_replicas entries indexed by replica id, or
 * NULL) receive each admitted replica's count and whether it met the
 * tolerance within max_iters. */
int sc_batch_minimize_until(sc_batch *b, int64_t max_iters, double dt, double max_step, double tolerance,
                            int64_t *iterations, int32_t *converged);
int sc_batch_get_state(sc_batch *b, int32_t id, double *r, double *v, double *F);
/* Overwrite r, v and/or F of a replica (NULL = keep); lists are not rebuilt. */
int sc_batch_set_state(sc_batch *b, int32_t id, const double *r, const double *v, const double *F);
/* Steps taken in the batch, rebuild count, and (if epot != NULL) the
 * potential energy of calculate_full_energy. */
int sc_batch_replica_stats(sc_batch *b, int32_t id, int64_t *steps, int64_t *rebuilds, double *epot);
/* The error state of one replica (SC_OK, SC_ERR_OVERFLOW or
 * SC_ERR_NONFINITE).  Replicas are independent: a failed replica can be
 * removed and the others continue unaffected. */
int sc_batch_replica_error(sc_batch *b, int32_t id, int32_t *code);
/* Short-list statistics of a replica (as sc_neighbor_stats). */
int sc_batch_neighbor_stats(sc_batch *b, int32_t id, int64_t *total_entries, int64_t *max_row);
/* Admitted replicas, their total particles and device bytes, and the device
 * time of the last sc_batch_step in milliseconds. */
int sc_batch_info(sc_batch *b, int32_t *active, int64_t *particles, int64_t *device_bytes, double *last_ms);

/* ---- launch shape and autotuner --------------------------------------
 * No reference counterpart: the reference's rebuild is CellListMap's CPU
 * cell list (src/NeighborLists.jl:46-52).  The per-slot neighbour-list build
 * kernel runs with block_size threads per block (64 or 256) and
 * lanes_per_particle lanes cooperating on one slot (1, 2, 4, 8, 16 or 32);
 * block_size 256 with lanes_per_particle 0 selects the per-slot radius
 * form (one lane per slot, searching only out to the slot's own largest
 * half-list radius).  A new context starts with a shape chosen by its
 * size: {64, 16} below 20000 particles, {256, 4} below 60000, {256, 1}
 * below 200000, else {256, 0}.
 * The lists, and so every result, are identical for every shape; only the
 * time differs (one lane per slot suits one large system, several lanes
 * suit batches of small replicas whose rebuild is latency-bound). */
typedef struct sc_tune_entry {
    int32_t block_size;
    int32_t lanes_per_particle;
    double ms_per_step; /* device time per step measured for this shape */
} sc_tune_entry;

#define SC_TUNE_SHAPES 13

int sc_set_build_launch(sc_ctx *ctx, int32_t block_size, int32_t lanes_per_particle);
int sc_batch_set_build_launch(sc_batch *b, int32_t block_size, int32_t lanes_per_particle);
/* Autotuner: for each of the SC_TUNE_SHAPES shapes, one untimed step and
 * then `steps` timed steps (real steps of the system: state, step and
 * rebuild counters advance, on-device noise only), then keeps the fastest
 * shape.  entries (may be NULL) receives up to max_entries measurements,
 * *best the index of the chosen one. */
int sc_autotune(sc_ctx *ctx, int64_t steps, sc_tune_entry *entries, int32_t max_entries, int32_t *best);
int sc_batch_autotune(sc_batch *b, int64_t steps, sc_tune_entry *entries, int32_t max_entries, int32_t *best);

#ifdef __cplusplus
}
#endif
#endif /* SC_H */
