// sc_internal.h -- device-side data layout of libsc (not part of the C-ABI).
//
// One "replica" is one independent periodic system (the reference's
// ArrayStruct + NeighborListStruct).  Every kernel runs over a table of
// tiles; a tile is SC_TB consecutive particle slots of one replica, so one
// launch can serve one system or many replicas of uneven N and mixed 2D/3D.
//
// HBM layout per replica (n particles, slots sorted by cell at each rebuild):
//   pos    double4[n]   {x, y, z, D}   (z = 0 in 2D); D rides with r because
//                                      every pair evaluation needs both
//   spos   uint4[n]     mixed precision: the same in 32-bit fixed point, the
//                                      16-byte record its pair loop gathers
//   v, F, disp          SoA doubles, component c at [c*n + i]
//   orig   int[n]       original (host) index of the particle in slot i
//   nn     int[n]       short-list length of slot i
//   nbr    int[cap*n]   short list, entry k of slot i at [k*n + i] so that a
//                       wavefront of consecutive slots reads consecutive words
//   cell_start int[ncell+1]  slot range of each cell (cell-sorted order),
//                       cell_start[ncell] = n
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SC_TB 256  // threads per tile / block (4 wavefronts)
#define SC_CHUNK_SHIFT 10
#define SC_CHUNK (1 << SC_CHUNK_SHIFT)  // cells per block of the two-level scan
#define SC_FUSED_CHUNKS 2048            // up to this many chunks the scan is one kernel
#define SC_CHUNK_SPLIT 8                // partial counters per chunk (spreads the binning atomics)
#define SC_WMAX 1024                    // particles of one LDS-staged build window
#define SC_GRID_CAP 2048                // blocks of the grid-stride rebuild kernels
#define SC_RUN_GRID_CAP 1024            // blocks of the LDS build
#define SC_WIN_MAX 2048                 // particles of one windowed-build LDS window
#define SC_WIN_CELLS 768                // cells of one window
#define SC_WIN_ROWS 64                  // own cell rows of one window run

// neighbour-list build kinds of a replica
enum { BUILD_SLOT = 0, BUILD_CELLS = 1, BUILD_WIN = 2, BUILD_ROWS = 3 };

#define REBUILD_KEEP_F 2

// error bits raised on the device
#define SC_EBIT_OVERFLOW 1
#define SC_EBIT_NONFINITE 2
#define SC_EBIT_NOISE 4  // a force evaluation found no injected noise block for its step

// last-block-done counters
enum { CTR_FORCE = 0, CTR_DRIFT, CTR_BUILD, CTR_ENERGY, CTR_EKIN, CTR_MISC, CTR_EKCOM, CTR_N };

struct RepDyn {
    unsigned long long maxdisp_bits;  // atomicMax over disp^2 (non-negative doubles)
    unsigned long long maxaux_bits;   // atomicMax scratch (minimiser step / force)
    int rebuild;                      // nonzero => the rebuild kernels run; bit 1 (REBUILD_KEEP_F):
                                      // requested between steps, the stored F is live and moves along
    int err;                          // SC_EBIT_*
    int pending;                      // 1 => vmean / vscale not yet applied to v
    int have_lists;
    unsigned int ctr[CTR_N];
    long long rebuilds;
    long long step;                   // force evaluations so far (noise counter)
    int min_state;                    // minimiser run to a tolerance: 0 running, 1 converged in
                                      // this iteration (its move still runs), 2 stopped; read
                                      // only while Rep::min_tol >= 0 (a tolerance run)
    int halt;                         // step chunk launched ahead: a step's drift fired the skin
                                      // trigger, so its force and every later kernel of the chunk
                                      // return at once until the host has run the rebuild
    long long min_iters;              // iterations done by the current tolerance run
    long long target;                 // steps launched ahead for a batch stop at this step count
                                      // (LLONG_MAX outside such a call)
    double vmean[3];
    double vscale;
    double Fmean[3];
    double result[4];                 // [0] energy, [1] kinetic, [2] minimiser scale, [3] max |F|
};

struct Rep {
    // ---- geometry / physics (immutable after creation)
    int dims, pot, sys, prec;
    int n;               // particles
    int nc;              // cells per side
    int ncell;           // nc^dims
    int srange;          // stencil half-width in cells (0 => single cell, all pairs)
    int cap;             // maxneighbours (row capacity)
    int rescale;
    int ntiles, tile0;   // tiles of this replica in the launch table
    int build_kind;      // BUILD_SLOT: per-slot sweep; BUILD_CELLS: LDS cell runs (stencil +-1);
                         // BUILD_WIN: LDS windows over cx x cy x cz own cells (any stencil);
                         // BUILD_ROWS: LDS-staged stencil rows, one wave per 64 slots
    int cx, cy, cz;      // own cells per build run along x, y, z
    int nruns, run0;     // build runs of this replica and their offset in the launch
    float wlo, whi;      // windowed build: FP32 d^2 below wlo is certainly <= R_v^2, above whi
                         // certainly > R_v^2; in between the FP64 test decides
    float hmarg;         // same margin for the half-list radius
    float dmax;          // largest diameter held so far (bounds every pair's half-list radius)
    float dmin;          // smallest diameter held so far
    // per-slot build with the fixed-point prefilter (bf_on): each candidate is
    // first tested on its 16-byte rebuild-time record bpos (fixed-point
    // coordinates, FP32 diameter) in FP32; d2 in fixed-point units^2 below
    // bf_lo is certainly <= R_v^2, above bf_hi certainly > R_v^2; the
    // half-list radius^2 rh2 (FP32, fixed-point units) is certainly passed
    // below rh2 * bf_hlo - bf_m and certainly failed above rh2 * bf_hhi + bf_m
    // (LJ: constant bounds bf_rh2lo / bf_rh2hi); in between the FP64 test on
    // pos_alt decides
    int bf_on;
    float bf_lo, bf_hi, bf_hlo, bf_hhi, bf_m, bf_rh2lo, bf_rh2hi;
    float bf_c1, bf_c2, bf_zeta;  // FP32 rcut S / 2, skin S (S = 2^32 / L), zeta
    double wcell;        // cell side L / nc
    double L, invL, cinv;          // box, 1/box, cells per unit length
    double fc2, rcut, skin, Rv2, thr;
    unsigned long long thr_bits;  // bit pattern of thr (= (skin/2)^2)
    double eps, lj_s2, lj_md, Uc;  // LJ: sigma^2 (specialised loops), find_mean_D = sigma^2
    double c0, c2, c4, zeta;
    double kBT, gamma, dt, m;
    double min_tol;      // minimiser: stop once max |F| <= min_tol (< 0: never)
    // mixed precision: each slot's position and diameter as 32-bit fixed point
    // (spos), u = x * uq mod 2^32 (uq = 2^32 / L, so a pair's minimum image is
    // the wrapping difference u_i - u_j read as a signed integer, times lq =
    // L / 2^32) and w = D * dqinv (dqinv a power of two, D = w * dq)
    double uq, lq, dq, dqinv;
    uint64_t seed;
    // ---- injected noise (NULL => Philox): block s at noise + (step - noise_step0) * n * dims,
    //      for 0 <= step - noise_step0 < noise_blocks
    const double *noise;
    long long noise_step0;
    long long noise_blocks;
    // ---- device buffers
    double4 *pos, *pos_alt;
    uint4 *spos;         // mixed precision only (else NULL): {ux, uy, uz, w} per slot, kept with pos
    uint4 *bpos;         // per-slot build prefilter (else NULL): {ux, uy, uz, FP32 bits of D} of pos_alt,
                         // written with the permutation at every rebuild
    double *v, *v_alt, *F, *disp;
    int *orig, *orig_alt;
    int *cell_of, *rank, *cell_cnt, *cell_start, *chunk_off, *key;
    int *chunk_cnt;      // particles per chunk of SC_CHUNK cells, SC_CHUNK_SPLIT partial counts
                         // per chunk [chunk * SC_CHUNK_SPLIT + part] (zero between rebuilds)
    int *nn, *nbr;
    double *part;        // per-tile partials, slot-major: part[slot * ntiles + tile], 8 slots
    double *stage;       // n*3 doubles staging for host transfers
    RepDyn *dyn;
};

struct Tile {
    int rep;
    int first;  // first slot of the tile within its replica
    int local;  // tile index within the replica
    int pad;
};

// launchers (sc_kernels.hip); every tile kernel is followed by the
// one-block-per-replica k_finish that reduces its tile partials
namespace sc {
int launch_kick_drift(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s);
int launch_copy(const double4 *src, double4 *dst, size_t n, hipStream_t s);
// ahead = true: launched as part of a chunk of steps enqueued ahead of the
// rebuild decision (RepDyn::halt)
int launch_force(const Tile *tiles, int ntiles, const Rep *reps, int nreps, int mode, int sysmask, hipStream_t s,
                 bool finish = true, bool ahead = false);
int launch_force_finish(const Rep *reps, int nreps, int mode, int sysmask, hipStream_t s, bool ahead = false);
int launch_kick_drift_only(const Tile *tiles, int ntiles, const Rep *reps, hipStream_t s, bool ahead = false);
int launch_drift_brownian(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s, bool ahead = false);
int launch_ekin_com(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s);
struct RebuildGeom {
    int max_chunks = 0;      // max over replicas of ceil(ncell / SC_CHUNK)
    int nruns = 0;           // LDS cell-run build runs over all replicas (BUILD_CELLS)
    int nwin = 0;            // windowed-build runs over all replicas (BUILD_WIN)
    int any_tile_build = 0;  // replicas use the per-slot build: bit 0 the FP64 sweep (k_build), bit 1 the
                             // fixed-point prefilter (k_build_fx)
    int any_rows_build = 0;  // replicas use the staged-row build: bit 2 some 2D replica, bit 3 some 3D one
    int *plan = nullptr;     // device work list, plan_ints(nreps, ntiles, nwin) ints (k_rebuild_plan)
    const int *flag = nullptr;  // one replica without a plan: its RepDyn::rebuild on the device
    int build_bs = 256;      // per-slot build launch shape: threads per block (64 or 256)
    int build_lpp = 0;       //   and lanes per slot (1 to 32, powers of two; 0 with 256 threads: the
                             //   per-slot radius form); the lists do not depend on it
};
bool build_shape_valid(int bs, int lpp);
// default per-slot build launch shape of a single system of n particles
inline void default_build_shape(int n, int &bs, int &lpp)
{
    if (n < 20000) {
        bs = 64;
        lpp = 16;
    } else if (n < 60000) {
        bs = 256;
        lpp = 4;
    } else if (n < 200000) {
        bs = 256;
        lpp = 1;
    } else {
        bs = 256;
        lpp = 0;
    }
}
// rebuild work list: header {tiles, replicas, window runs, 0}, then the
// replicas, the tiles and the (replica, run) pairs of the rebuilding replicas
inline int plan_ints(int nreps, int ntiles, int nwin) { return 4 + nreps + ntiles + 2 * nwin; }
// the per-slot build launch shapes tried by the autotuner, {block, lanes};
// lanes 0: the per-slot radius form (k_build_px)
constexpr int kBuildShapes[13][2] = {{256, 1}, {256, 2}, {256, 4}, {256, 8}, {256, 16}, {256, 32}, {64, 1},
                                     {64, 2},  {64, 4},  {64, 8},  {64, 16}, {64, 32},  {256, 0}};
// bookkeep=false when the next launch is a velocity-Verlet force finish,
// which does the rebuild bookkeeping itself
int launch_rebuild(const Tile *tiles, int ntiles, const Rep *reps, int nreps, const RebuildGeom &g, hipStream_t s,
                   bool bookkeep = true);
int launch_request_rebuild(const Rep *reps, int nreps, hipStream_t s);
// rebuild bookkeeping alone (replicas of the dynamics in sysmask: bit 0 Newtonian, 1 Langevin, 2 Brownian)
int launch_rebuild_done(const Rep *reps, int nreps, hipStream_t s, int sysmask = 7);
// per replica: what = 0 its error bits, 1 its minimiser stop state, 2 its
// chunk state (bit 0 halt, bit 1 rebuild flag, bits 2.. steps left to its target)
int launch_gather_err(const Rep *reps, int nreps, int *out, hipStream_t s, int what = 0);
// step targets of a batch call launched ahead: target = step + nsteps (nsteps >= 0), or none (< 0)
int launch_set_target(const Rep *reps, int nreps, long long nsteps, hipStream_t s);
// after a chunk launched ahead: the force (mode 1) and finish of the step whose
// drift halted each halted velocity-Verlet replica, for those replicas only
int launch_force_resume(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s);
int launch_finalize(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s);
// per (optional): E_i / 2 of every particle at its original index
int launch_energy(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s, double *per = nullptr);
int launch_energy_nn(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s);
int launch_ekin(const Tile *tiles, int ntiles, const Rep *reps, int nreps, int apply_rescale, hipStream_t s);
int launch_set_rold(const Tile *tiles, int ntiles, const Rep *reps, double *rold, hipStream_t s);
int launch_f2s(const Tile *tiles, int ntiles, const Rep *reps, int nreps, const double *rold, double kx,
               hipStream_t s);
constexpr int QL_LMAX = 8;  // largest Steinhardt order l served
int launch_steinhardt(const Tile *tiles, int ntiles, const Rep *reps, int nreps, int l, double qc2, int averaged,
                      double *qlm, int *nq, hipStream_t s);
int launch_from_host_order(const Tile *tiles, int ntiles, const Rep *reps, int field, hipStream_t s);
int launch_to_host_order(const Tile *tiles, int ntiles, const Rep *reps, int field, hipStream_t s);
int launch_full_rows(const Tile *tiles, int ntiles, const Rep *reps, int *rows, int *counts, int cap, hipStream_t s);
// reference-layout Int64 rows (1-based, zero past each count) and counts of
// the full (half = 0) or half (half = 1) list, written into caller memory;
// *flag (cleared first) becomes 1 if a row is longer than maxneigh
int launch_ref_rows(const Tile *tiles, int ntiles, const Rep *reps, int half, long long *rows, long long *counts,
                    int maxneigh, int *flag, hipStream_t s);
int launch_min_scale(const Tile *tiles, int ntiles, const Rep *reps, int nreps, double dt, double max_step,
                     hipStream_t s);
int launch_min_move(const Tile *tiles, int ntiles, const Rep *reps, int nreps, double dt, hipStream_t s);
// tolerance-run markers: begin resets the stop state and iteration count,
// end unblocks the force kernel
int launch_min_mark(const Rep *reps, int nreps, bool begin, hipStream_t s);
}  // namespace sc

// replica setup (sc_replica.hip)
#include <string>
#include <vector>
struct sc_params;
namespace sc {
bool params_valid(const sc_params *p);
bool diameters_valid(const double *D, int64_t n);  // every D[i] finite and > 0
int64_t replica_bytes(const sc_params *p, const double *D);
int replica_init(Rep &h, const sc_params *p, const double *D, std::vector<void *> &allocs, std::string &err);
// fixed-point diameter scale of the mixed form for diameters up to maxD, and
// one diameter in it (the device's rint(D * dqinv), bit for bit)
void set_diameter_scale(Rep &h, double maxD);
// FP32 margins of the per-slot build's fixed-point prefilter from the current
// diameter range; clears bf_on where they do not hold (zeta < 0 or
// zeta * (max D - min D) > 1/2 for the polydisperse potential)
void build_prefilter_setup(Rep &h);
unsigned diameter_fixed(double D, double dqinv);
void free_all(std::vector<void *> &allocs);
}  // namespace sc

// fields for the host-order transfer kernels
enum { FLD_R = 0, FLD_V, FLD_F, FLD_D, FLD_DISP, FLD_DISP2 };
