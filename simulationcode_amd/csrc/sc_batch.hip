// sc_batch.hip -- many independent replicas advanced by shared launches.
//
// Every replica keeps its own device state (sc_replica.hip); the batch owns
// one compact descriptor array and one tile table over all active replicas,
// so each kernel of a step is launched once for the whole batch whatever the
// replicas' N, dimension, potential, dynamics or precision.  A step is
//   [Brownian: forces (+noise), mean force, drift]  [VV: first kick + drift]
//   [rebuild where a replica's skin trigger fired]  [VV: forces, second kick,
//   COM]  [rescale where requested]
// and may be captured once into a hipGraph and replayed.  Replicas use the
// on-device Philox noise keyed by their own seed, so a replica's trajectory is
// bitwise the same alone and in any mix.
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sc.h"
#include "sc_internal.h"

namespace {

struct Slot {
    bool used = false;
    Rep h{};
    sc_params p{};
    std::vector<void *> allocs;
    long long steps = 0;
    long long bytes = 0;
};

}  // namespace

struct sc_batch {
    int max_reps = 0;
    std::vector<Slot> slots;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Rep *d_reps = nullptr;
    Tile *d_tiles = nullptr;
    int tiles_cap = 0, ntiles = 0, nact = 0;
    std::vector<int> act;  // slot of each compact replica index
    sc::RebuildGeom geom{};
    int *d_plan = nullptr;
    int plan_cap = 0;
    int *d_err = nullptr;  // error bits of the active replicas (max_reps ints)
    int build_bs = 256, build_lpp = 4;  // per-slot build launch shape (small replicas: several lanes)
    bool dirty = true, has_vv = false, has_bd = false, has_rescale = false;
    bool use_graph = false;
    int64_t pending = -1;  // steps enqueued by sc_batch_step_async and not yet waited for
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::string err;
    double last_ms = 0.0;
    int device = 0;
    int chunk = 8;                // steps launched ahead per chunk (sc_batch_set_ahead; 0: eager)
    int *h_state = nullptr;       // pinned: per-replica chunk state (k_gather_err what = 2)
    std::thread worker;           // sc_batch_step_async's chunk loop, joined by sc_batch_wait
    int worker_rc = SC_OK;
};

namespace {

int bfail(sc_batch *b, int code, const std::string &m)
{
    if (b) b->err = m;
    return code;
}

#define BCHK(b, expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return bfail(b, SC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BK(b, expr)                                                                              \
    do {                                                                                         \
        if ((expr) != 0) return bfail(b, SC_ERR_HIP, std::string("kernel launch failed: ") +   \
                                                         hipGetErrorString(hipGetLastError())); \
    } while (0)

void drop_graph(sc_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// Rebuild the compact descriptor array and the tile table after admissions
// or retirements.
int sync_tables(sc_batch *b)
{
    if (b->pending >= 0) {  // settle an asynchronous step first
        int rc = sc_batch_wait(b);
        if (rc) return rc;
    }
    if (!b->dirty) return SC_OK;
    BCHK(b, hipStreamSynchronize(b->s));
    drop_graph(b);
    b->act.clear();
    for (int i = 0; i < (int)b->slots.size(); i++)
        if (b->slots[i].used) b->act.push_back(i);
    b->nact = (int)b->act.size();
    std::vector<Rep> reps;
    std::vector<Tile> tiles;
    b->geom = sc::RebuildGeom{};
    b->has_vv = b->has_bd = b->has_rescale = false;
    int run0 = 0;
    for (int r = 0; r < b->nact; r++) {
        Slot &sl = b->slots[b->act[r]];
        sl.h.tile0 = (int)tiles.size();
        sl.h.run0 = run0;
        if (sl.h.build_kind == BUILD_CELLS) run0 += sl.h.nruns;
        if (sl.h.build_kind == BUILD_WIN) b->geom.nwin += sl.h.nruns;
        for (int t = 0; t < sl.h.ntiles; t++) tiles.push_back(Tile{r, t * SC_TB, t, 0});
        reps.push_back(sl.h);
        b->geom.max_chunks = std::max(b->geom.max_chunks, (sl.h.ncell + SC_CHUNK - 1) / SC_CHUNK);
        if (sl.h.build_kind == BUILD_SLOT) b->geom.any_tile_build |= sl.h.bf_on ? 2 : 1;
        if (sl.h.build_kind == BUILD_ROWS) b->geom.any_rows_build |= 1 << sl.h.dims;
        if (sl.h.sys == SC_SYS_BROWNIAN)
            b->has_bd = true;
        else
            b->has_vv = true;
        if (sl.h.rescale) b->has_rescale = true;
    }
    b->geom.nruns = run0;
    b->ntiles = (int)tiles.size();
    const int plan_need = sc::plan_ints(b->nact, b->ntiles, b->geom.nwin);
    if (plan_need > b->plan_cap) {
        const int cap = std::max(plan_need, 2 * b->plan_cap);
        if (b->d_plan) (void)hipFree(b->d_plan);
        b->d_plan = nullptr;
        b->plan_cap = 0;
        if (hipMalloc(&b->d_plan, sizeof(int) * cap) != hipSuccess) return bfail(b, SC_ERR_OOM, "rebuild plan");
        b->plan_cap = cap;
    }
    b->geom.plan = b->d_plan;
    b->geom.build_bs = b->build_bs;
    b->geom.build_lpp = b->build_lpp;
    if (b->ntiles > b->tiles_cap) {
        if (b->d_tiles) (void)hipFree(b->d_tiles);
        b->d_tiles = nullptr;
        int cap = std::max(b->ntiles, 2 * b->tiles_cap);
        if (hipMalloc(&b->d_tiles, sizeof(Tile) * cap) != hipSuccess) return bfail(b, SC_ERR_OOM, "tile table");
        b->tiles_cap = cap;
    }
    if (b->ntiles) BCHK(b, hipMemcpy(b->d_tiles, tiles.data(), sizeof(Tile) * b->ntiles, hipMemcpyHostToDevice));
    if (b->nact) BCHK(b, hipMemcpy(b->d_reps, reps.data(), sizeof(Rep) * b->nact, hipMemcpyHostToDevice));
    b->dirty = false;
    return SC_OK;
}

int enqueue_step(sc_batch *b)
{
    const int nt = b->ntiles, nr = b->nact;
    if (b->has_bd) {
        BK(b, sc::launch_force(b->d_tiles, nt, b->d_reps, nr, 2, 4, b->s));
        BK(b, sc::launch_drift_brownian(b->d_tiles, nt, b->d_reps, nr, b->s));
    }
    if (b->has_vv) BK(b, sc::launch_kick_drift(b->d_tiles, nt, b->d_reps, nr, b->s));
    // an all-Brownian batch leaves the rebuild bookkeeping to the next step's
    // force finish (and the end of the call): one launch less per step
    BK(b, sc::launch_rebuild(b->d_tiles, nt, b->d_reps, nr, b->geom, b->s, b->has_vv));
    if (b->has_vv) BK(b, sc::launch_force(b->d_tiles, nt, b->d_reps, nr, 1, 3, b->s));
    if (b->has_rescale) BK(b, sc::launch_ekin_com(b->d_tiles, nt, b->d_reps, nr, b->s));
    return SC_OK;
}

// Steps enqueued ahead in chunks, as a single system does (sc_api.hip), for
// every replica at once.  Within a chunk each replica runs its own steps
// until its drift fires its skin trigger (the replica then halts: its later
// kernels of the chunk return at once, RepDyn::halt) or it reaches this
// call's step target; the others go on.  After each chunk one copy reads
// every replica's state, and the halted replicas are rebuilt together in one
// rebuild chain (velocity Verlet: followed by the force and finish of the
// halted step, for those replicas only), then the next chunk continues.
// Each replica's kernels and their order are those of enqueue_step, so its
// trajectory is bitwise the same; the rebuild work of a chunk is shared by
// all replicas that triggered in it instead of a rebuild chain per step.
int steps_ahead_batch(sc_batch *b, int64_t nsteps)
{
    const int nt = b->ntiles, nr = b->nact;
    BK(b, sc::launch_set_target(b->d_reps, nr, nsteps, b->s));
    int64_t left = nsteps;
    while (left > 0) {
        const int64_t m = std::min<int64_t>(b->chunk, left);
        for (int64_t i = 0; i < m; i++) {
            if (b->has_bd) {
                BK(b, sc::launch_force(b->d_tiles, nt, b->d_reps, nr, 2, 4, b->s, false, true));
                BK(b, sc::launch_force_finish(b->d_reps, nr, 2, 4, b->s, true));
                BK(b, sc::launch_drift_brownian(b->d_tiles, nt, b->d_reps, nr, b->s, true));
            }
            if (b->has_vv) {
                BK(b, sc::launch_kick_drift_only(b->d_tiles, nt, b->d_reps, b->s, true));
                BK(b, sc::launch_force(b->d_tiles, nt, b->d_reps, nr, 1, 3, b->s, false, true));
                BK(b, sc::launch_force_finish(b->d_reps, nr, 1, 3, b->s, true));
            }
        }
        BK(b, sc::launch_gather_err(b->d_reps, nr, b->d_err, b->s, 2));
        BCHK(b, hipMemcpyAsync(b->h_state, b->d_err, sizeof(int) * nr, hipMemcpyDeviceToHost, b->s));
        BCHK(b, hipStreamSynchronize(b->s));
        bool halted = false;
        int64_t most = 0;
        for (int r = 0; r < nr; r++) {
            const int st = b->h_state[r];
            int64_t l = st >> 2;
            if (st & 3) {
                halted = true;
                // a halted velocity-Verlet replica completes its halted step below
                if (b->slots[b->act[r]].h.sys != SC_SYS_BROWNIAN && (st & 1)) l -= 1;
            }
            most = std::max(most, l);
        }
        if (halted) {
            BK(b, sc::launch_rebuild(b->d_tiles, nt, b->d_reps, nr, b->geom, b->s, false));
            if (b->has_vv) BK(b, sc::launch_force_resume(b->d_tiles, nt, b->d_reps, nr, b->s));
            if (b->has_bd) BK(b, sc::launch_rebuild_done(b->d_reps, nr, b->s, 4));
        }
        // every replica short of its target advances at least one step per
        // chunk (a Brownian replica halts only after the step whose drift
        // fired, a velocity-Verlet one completes that step above)
        if (most >= left) {
            BK(b, sc::launch_set_target(b->d_reps, nr, -1, b->s));
            return bfail(b, SC_ERR_STATE, "chunked steps made no progress");
        }
        left = most;
    }
    BK(b, sc::launch_set_target(b->d_reps, nr, -1, b->s));
    return SC_OK;
}

int check_slot(sc_batch *b, int id)
{
    if (!b || id < 0 || id >= (int)b->slots.size() || !b->slots[id].used) return SC_ERR_ARG;
    return sync_tables(b);
}

int replica_index(sc_batch *b, int id)
{
    for (int r = 0; r < b->nact; r++)
        if (b->act[r] == id) return r;
    return -1;
}

int batch_errors(sc_batch *b)
{
    if (b->nact == 0) return SC_OK;
    std::vector<int> err(b->nact);
    BK(b, sc::launch_gather_err(b->d_reps, b->nact, b->d_err, b->s));
    BCHK(b, hipMemcpyAsync(err.data(), b->d_err, sizeof(int) * b->nact, hipMemcpyDeviceToHost, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    for (int r = 0; r < b->nact; r++) {
        if (err[r] & SC_EBIT_OVERFLOW)
            return bfail(b, SC_ERR_OVERFLOW, "replica " + std::to_string(b->act[r]) + ": neighbour list overflow");
        if (err[r] & SC_EBIT_NONFINITE)
            return bfail(b, SC_ERR_NONFINITE, "replica " + std::to_string(b->act[r]) + ": non-finite position");
    }
    return SC_OK;
}

}  // namespace

extern "C" {

int sc_batch_create(sc_batch **out, int32_t max_replicas)
{
    if (!out || max_replicas < 1 || max_replicas > 65536) return SC_ERR_ARG;
    sc_batch *b = new sc_batch();
    *out = b;
    b->max_reps = max_replicas;
    b->slots.resize(max_replicas);
    if (hipStreamCreate(&b->s) != hipSuccess) return bfail(b, SC_ERR_HIP, "hipStreamCreate failed");
    BCHK(b, hipGetDevice(&b->device));
    if (hipHostMalloc((void **)&b->h_state, sizeof(int) * max_replicas, hipHostMallocDefault) != hipSuccess) {
        b->h_state = nullptr;
        return bfail(b, SC_ERR_OOM, "pinned chunk-state buffer");
    }
    BCHK(b, hipEventCreate(&b->ev0));
    BCHK(b, hipEventCreate(&b->ev1));
    if (hipMalloc(&b->d_reps, sizeof(Rep) * max_replicas) != hipSuccess)
        return bfail(b, SC_ERR_OOM, "replica descriptor table");
    if (hipMalloc(&b->d_err, sizeof(int) * max_replicas) != hipSuccess)
        return bfail(b, SC_ERR_OOM, "replica error array");
    return SC_OK;
}

void sc_batch_destroy(sc_batch *b)
{
    if (!b) return;
    if (b->worker.joinable()) b->worker.join();
    if (b->s) (void)hipStreamSynchronize(b->s);
    if (b->h_state) (void)hipHostFree(b->h_state);
    drop_graph(b);
    for (auto &sl : b->slots) sc::free_all(sl.allocs);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_plan) (void)hipFree(b->d_plan);
    if (b->d_err) (void)hipFree(b->d_err);
    if (b->d_reps) (void)hipFree(b->d_reps);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
}

const char *sc_batch_last_error(const sc_batch *b) { return b ? b->err.c_str() : "null batch"; }

int64_t sc_replica_device_bytes(const sc_params *p, const double *D)
{
    if (!p || !D || !sc::params_valid(p) || !sc::diameters_valid(D, p->N)) return -1;
    return sc::replica_bytes(p, D);
}

int sc_batch_add(sc_batch *b, const sc_params *p, const double *D, const double *r, const double *v, int32_t *id)
{
    if (!b || !p || !D || !r || !id) return SC_ERR_ARG;
    if (!sc::params_valid(p)) return bfail(b, SC_ERR_ARG, "invalid parameters");
    if (!sc::diameters_valid(D, p->N)) return bfail(b, SC_ERR_ARG, "diameters must be finite and positive");
    if (b->pending >= 0) {  // an asynchronous step's worker reads the slots: settle it first
        int rc = sc_batch_wait(b);
        if (rc) return rc;
    }
    int slot = -1;
    for (int i = 0; i < b->max_reps; i++)
        if (!b->slots[i].used) {
            slot = i;
            break;
        }
    if (slot < 0) return bfail(b, SC_ERR_OOM, "no free replica slot");
    Slot &sl = b->slots[slot];
    sl.allocs.clear();
    int rc = sc::replica_init(sl.h, p, D, sl.allocs, b->err);
    if (rc != SC_OK) {
        sc::free_all(sl.allocs);
        return rc;
    }
    sl.p = *p;
    sl.used = true;
    sl.steps = 0;
    sl.bytes = sc::replica_bytes(p, D);
    b->dirty = true;
    if ((rc = sync_tables(b))) return rc;
    const Rep &h = b->slots[slot].h;
    const size_t nd = (size_t)h.n * h.dims;
    BCHK(b, hipMemcpyAsync(h.stage, r, nd * sizeof(double), hipMemcpyHostToDevice, b->s));
    BK(b, sc::launch_from_host_order(b->d_tiles + h.tile0, h.ntiles, b->d_reps, FLD_R, b->s));
    if (v) {
        BCHK(b, hipMemcpyAsync(h.stage, v, nd * sizeof(double), hipMemcpyHostToDevice, b->s));
        BK(b, sc::launch_from_host_order(b->d_tiles + h.tile0, h.ntiles, b->d_reps, FLD_V, b->s));
    }
    int one = 1;
    BCHK(b, hipMemcpyAsync(&h.dyn->rebuild, &one, sizeof(int), hipMemcpyHostToDevice, b->s));
    BK(b, sc::launch_rebuild(b->d_tiles, b->ntiles, b->d_reps, b->nact, b->geom, b->s, true));
    RepDyn d;
    BCHK(b, hipMemcpyAsync(&d, h.dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    // only this replica's state decides the admission (the others' errors
    // surface at their next step); a replica that fails its first rebuild is
    // not admitted and its slot is freed again
    if (d.err) {
        const int code = (d.err & SC_EBIT_OVERFLOW) ? SC_ERR_OVERFLOW : SC_ERR_NONFINITE;
        int rc2 = sc_batch_remove(b, slot);
        if (rc2) return rc2;
        return bfail(b, code, code == SC_ERR_OVERFLOW ? "admission: neighbour list overflow (maxneighbours)"
                                                      : "admission: non-finite state");
    }
    *id = slot;
    return SC_OK;
}

int sc_batch_remove(sc_batch *b, int32_t id)
{
    if (b && b->pending >= 0) {
        int rc = sc_batch_wait(b);
        if (rc) return rc;
    }
    if (!b || id < 0 || id >= (int)b->slots.size() || !b->slots[id].used) return SC_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->s));
    sc::free_all(b->slots[id].allocs);
    b->slots[id].used = false;
    b->dirty = true;
    return SC_OK;
}

int sc_batch_set_ahead(sc_batch *b, int32_t chunk)
{
    if (b && b->pending >= 0) {
        int rc = sc_batch_wait(b);
        if (rc) return rc;
    }
    if (!b || chunk < 0 || chunk > 4096) return SC_ERR_ARG;
    b->chunk = chunk;
    return SC_OK;
}

int sc_batch_set_graph(sc_batch *b, int32_t on)
{
    if (b && b->pending >= 0) {
        int rc = sc_batch_wait(b);
        if (rc) return rc;
    }
    if (!b) return SC_ERR_ARG;
    b->use_graph = on != 0;
    return SC_OK;
}

int sc_batch_step_async(sc_batch *b, int64_t nsteps)
{
    if (!b || nsteps < 0) return SC_ERR_ARG;
    if (b->pending >= 0) return bfail(b, SC_ERR_STATE, "a step is already in flight (call sc_batch_wait)");
    int rc = sync_tables(b);
    if (rc) return rc;
    if (b->nact == 0 || nsteps == 0) {
        b->pending = 0;
        b->last_ms = 0.0;
        return SC_OK;
    }
    BCHK(b, hipEventRecord(b->ev0, b->s));
    if (b->chunk > 0 && !b->use_graph && !b->has_rescale) {
        // the chunk loop polls the device between chunks: it runs on a worker
        // thread (on the batch's device and stream) so this call returns at once
        b->worker_rc = SC_OK;
        b->pending = nsteps;
        b->worker = std::thread([b, nsteps]() {
            int rc2 = hipSetDevice(b->device) == hipSuccess ? SC_OK : bfail(b, SC_ERR_HIP, "hipSetDevice failed");
            if (!rc2) rc2 = steps_ahead_batch(b, nsteps);
            if (!rc2 && sc::launch_finalize(b->d_tiles, b->ntiles, b->d_reps, b->nact, b->s))
                rc2 = bfail(b, SC_ERR_HIP, "kernel launch failed (finalize)");
            if (!rc2 && hipEventRecord(b->ev1, b->s) != hipSuccess) rc2 = bfail(b, SC_ERR_HIP, "hipEventRecord failed");
            b->worker_rc = rc2;
        });
        return SC_OK;
    }
    if (b->use_graph) {
        if (!b->gexec) {
            BCHK(b, hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
            rc = enqueue_step(b);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->s, &g);
            if (rc) return rc;
            if (e != hipSuccess) return bfail(b, SC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            b->graph = g;
            BCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
        }
        for (int64_t s = 0; s < nsteps; s++) BCHK(b, hipGraphLaunch(b->gexec, b->s));
    } else {
        for (int64_t s = 0; s < nsteps; s++)
            if ((rc = enqueue_step(b))) return rc;
    }
    if (!b->has_vv) BK(b, sc::launch_rebuild_done(b->d_reps, b->nact, b->s));
    BK(b, sc::launch_finalize(b->d_tiles, b->ntiles, b->d_reps, b->nact, b->s));
    BCHK(b, hipEventRecord(b->ev1, b->s));
    b->pending = nsteps;
    return SC_OK;
}

int sc_batch_wait(sc_batch *b)
{
    if (!b) return SC_ERR_ARG;
    if (b->pending < 0) return SC_OK;
    const int64_t n = b->pending;
    b->pending = -1;
    if (b->worker.joinable()) {
        b->worker.join();
        if (b->worker_rc) return b->worker_rc;
    }
    if (n == 0) return SC_OK;
    BCHK(b, hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    for (int r = 0; r < b->nact; r++) b->slots[b->act[r]].steps += n;
    return batch_errors(b);
}

int sc_batch_step(sc_batch *b, int64_t nsteps)
{
    int rc = sc_batch_step_async(b, nsteps);
    if (rc) return rc;
    return sc_batch_wait(b);
}

int sc_batch_minimize(sc_batch *b, int64_t iters, double dt, double max_step)
{
    if (!b || iters < 0 || !(dt > 0.0) || !(max_step > 0.0)) return SC_ERR_ARG;
    int rc = sync_tables(b);
    if (rc) return rc;
    if (b->nact == 0) return SC_OK;
    // pair forces only: Newtonian views of the descriptors for the descent
    std::vector<Rep> view(b->nact);
    for (int r = 0; r < b->nact; r++) {
        view[r] = b->slots[b->act[r]].h;
        view[r].sys = SC_SYS_NEWTONIAN;
    }
    BCHK(b, hipMemcpyAsync(b->d_reps, view.data(), sizeof(Rep) * b->nact, hipMemcpyHostToDevice, b->s));
    const int nt = b->ntiles, nr = b->nact;
    for (int64_t it = 0; it < iters; it++) {
        BK(b, sc::launch_force(b->d_tiles, nt, b->d_reps, nr, 0, 7, b->s));
        BK(b, sc::launch_min_scale(b->d_tiles, nt, b->d_reps, nr, dt, max_step, b->s));
        BK(b, sc::launch_min_move(b->d_tiles, nt, b->d_reps, nr, dt, b->s));
        BK(b, sc::launch_rebuild(b->d_tiles, nt, b->d_reps, nr, b->geom, b->s, true));
    }
    for (int r = 0; r < b->nact; r++) view[r] = b->slots[b->act[r]].h;
    BCHK(b, hipMemcpyAsync(b->d_reps, view.data(), sizeof(Rep) * b->nact, hipMemcpyHostToDevice, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    return batch_errors(b);
}

int sc_batch_minimize_until(sc_batch *b, int64_t max_iters, double dt, double max_step, double tolerance,
                            int64_t *iterations, int32_t *converged)
{
    if (!b || max_iters < 0 || !(dt > 0.0) || !(max_step > 0.0) || !(tolerance >= 0.0)) return SC_ERR_ARG;
    int rc = sync_tables(b);
    if (rc) return rc;
    if (b->nact == 0) return SC_OK;
    // the descriptors are re-uploaded from the slots by the next call if this
    // one stops early (views left on the device)
    b->dirty = true;
    const int nt = b->ntiles, nr = b->nact;
    std::vector<Rep> view(nr);
    for (int r = 0; r < nr; r++) {
        view[r] = b->slots[b->act[r]].h;
        view[r].sys = SC_SYS_NEWTONIAN;
        view[r].min_tol = tolerance;
    }
    BCHK(b, hipMemcpyAsync(b->d_reps, view.data(), sizeof(Rep) * nr, hipMemcpyHostToDevice, b->s));
    BK(b, sc::launch_min_mark(b->d_reps, nr, true, b->s));
    // every replica stops on its own (its kernels return once it has met the
    // tolerance); the host polls the stop states between growing chunks
    std::vector<int> st(nr, 0), er(nr, 0);
    int64_t done = 0, chunk = 16;
    while (done < max_iters) {
        const int64_t k = std::min(chunk, max_iters - done);
        for (int64_t it = 0; it < k; it++) {
            BK(b, sc::launch_force(b->d_tiles, nt, b->d_reps, nr, 0, 7, b->s));
            BK(b, sc::launch_min_scale(b->d_tiles, nt, b->d_reps, nr, dt, max_step, b->s));
            BK(b, sc::launch_min_move(b->d_tiles, nt, b->d_reps, nr, dt, b->s));
            BK(b, sc::launch_rebuild(b->d_tiles, nt, b->d_reps, nr, b->geom, b->s, true));
        }
        done += k;
        BK(b, sc::launch_gather_err(b->d_reps, nr, b->d_err, b->s, 1));
        BCHK(b, hipMemcpyAsync(st.data(), b->d_err, sizeof(int) * nr, hipMemcpyDeviceToHost, b->s));
        BK(b, sc::launch_gather_err(b->d_reps, nr, b->d_err, b->s, 0));
        BCHK(b, hipMemcpyAsync(er.data(), b->d_err, sizeof(int) * nr, hipMemcpyDeviceToHost, b->s));
        BCHK(b, hipStreamSynchronize(b->s));
        bool all = true;
        for (int r = 0; r < nr; r++) all = all && (st[r] != 0 || er[r] != 0);
        if (all) break;
        chunk = std::min<int64_t>(chunk * 2, 1024);
    }
    BK(b, sc::launch_min_mark(b->d_reps, nr, false, b->s));
    std::vector<RepDyn> dyn(nr);
    for (int r = 0; r < nr; r++) {
        view[r] = b->slots[b->act[r]].h;
        BCHK(b, hipMemcpyAsync(&dyn[r], view[r].dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, b->s));
    }
    BCHK(b, hipMemcpyAsync(b->d_reps, view.data(), sizeof(Rep) * nr, hipMemcpyHostToDevice, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    b->dirty = false;
    for (int r = 0; r < nr; r++) {
        if (iterations) iterations[b->act[r]] = dyn[r].min_iters;
        if (converged) converged[b->act[r]] = st[r] != 0 ? 1 : 0;
    }
    return batch_errors(b);
}

int sc_batch_get_state(sc_batch *b, int32_t id, double *r, double *v, double *F)
{
    int rc = check_slot(b, id);
    if (rc) return rc;
    const Rep &h = b->slots[id].h;
    const size_t nd = (size_t)h.n * h.dims;
    double *dst[3] = {r, v, F};
    const int fld[3] = {FLD_R, FLD_V, FLD_F};
    for (int q = 0; q < 3; q++) {
        if (!dst[q]) continue;
        BK(b, sc::launch_to_host_order(b->d_tiles + h.tile0, h.ntiles, b->d_reps, fld[q], b->s));
        BCHK(b, hipMemcpyAsync(dst[q], h.stage, nd * sizeof(double), hipMemcpyDeviceToHost, b->s));
        BCHK(b, hipStreamSynchronize(b->s));
    }
    return SC_OK;
}

int sc_batch_set_state(sc_batch *b, int32_t id, const double *r, const double *v, const double *F)
{
    int rc = check_slot(b, id);
    if (rc) return rc;
    const Rep &h = b->slots[id].h;
    const size_t nd = (size_t)h.n * h.dims;
    const double *src[3] = {r, v, F};
    const int fld[3] = {FLD_R, FLD_V, FLD_F};
    for (int q = 0; q < 3; q++) {
        if (!src[q]) continue;
        BCHK(b, hipMemcpyAsync(h.stage, src[q], nd * sizeof(double), hipMemcpyHostToDevice, b->s));
        BK(b, sc::launch_from_host_order(b->d_tiles + h.tile0, h.ntiles, b->d_reps, fld[q], b->s));
    }
    BCHK(b, hipStreamSynchronize(b->s));
    return SC_OK;
}

int sc_batch_replica_stats(sc_batch *b, int32_t id, int64_t *steps, int64_t *rebuilds, double *epot)
{
    int rc = check_slot(b, id);
    if (rc) return rc;
    const Rep &h = b->slots[id].h;
    if (epot) {
        const int ri = replica_index(b, id);
        if (ri < 0) return SC_ERR_ARG;
        BK(b, sc::launch_energy(b->d_tiles + h.tile0, h.ntiles, b->d_reps, b->nact, b->s));
    }
    RepDyn d;
    BCHK(b, hipMemcpyAsync(&d, h.dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    if (steps) *steps = b->slots[id].steps;
    if (rebuilds) *rebuilds = d.rebuilds;
    if (epot) *epot = d.result[0];
    return SC_OK;
}

int sc_batch_neighbor_stats(sc_batch *b, int32_t id, int64_t *total_entries, int64_t *max_row)
{
    int rc = check_slot(b, id);
    if (rc) return rc;
    const Rep &h = b->slots[id].h;
    std::vector<int> nn(h.n);
    BCHK(b, hipMemcpyAsync(nn.data(), h.nn, sizeof(int) * h.n, hipMemcpyDeviceToHost, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    int64_t t = 0, m = 0;
    for (int x : nn) {
        t += x;
        m = std::max<int64_t>(m, x);
    }
    if (total_entries) *total_entries = t;
    if (max_row) *max_row = m;
    return SC_OK;
}

int sc_batch_replica_error(sc_batch *b, int32_t id, int32_t *code)
{
    int rc = check_slot(b, id);
    if (rc) return rc;
    RepDyn d;
    BCHK(b, hipMemcpyAsync(&d, b->slots[id].h.dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, b->s));
    BCHK(b, hipStreamSynchronize(b->s));
    *code = (d.err & SC_EBIT_OVERFLOW) ? SC_ERR_OVERFLOW : ((d.err & SC_EBIT_NONFINITE) ? SC_ERR_NONFINITE : SC_OK);
    return SC_OK;
}

int sc_batch_info(sc_batch *b, int32_t *active, int64_t *particles, int64_t *device_bytes, double *last_ms)
{
    if (!b) return SC_ERR_ARG;
    int rc = sync_tables(b);
    if (rc) return rc;
    int64_t np = 0, by = 0;
    for (int r = 0; r < b->nact; r++) {
        np += b->slots[b->act[r]].h.n;
        by += b->slots[b->act[r]].bytes;
    }
    if (active) *active = b->nact;
    if (particles) *particles = np;
    if (device_bytes) *device_bytes = by;
    if (last_ms) *last_ms = b->last_ms;
    return SC_OK;
}

int sc_batch_set_build_launch(sc_batch *b, int32_t block_size, int32_t lanes_per_particle)
{
    if (b && b->pending >= 0) {
        int rc = sc_batch_wait(b);
        if (rc) return rc;
    }
    if (!b || !sc::build_shape_valid(block_size, lanes_per_particle)) return SC_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->s));
    b->build_bs = b->geom.build_bs = block_size;
    b->build_lpp = b->geom.build_lpp = lanes_per_particle;
    drop_graph(b);  // a captured step holds the old launch
    return SC_OK;
}

int sc_batch_autotune(sc_batch *b, int64_t steps, sc_tune_entry *entries, int32_t max_entries, int32_t *best)
{
    if (!b || steps < 1 || max_entries < 0) return SC_ERR_ARG;
    int rc = sync_tables(b);
    if (rc) return rc;
    if (b->nact == 0) return bfail(b, SC_ERR_STATE, "autotune needs admitted replicas");
    int bi = 0;
    double bms = 0.0;
    for (int q = 0; q < SC_TUNE_SHAPES; q++) {
        if ((rc = sc_batch_set_build_launch(b, sc::kBuildShapes[q][0], sc::kBuildShapes[q][1]))) return rc;
        if ((rc = sc_batch_step(b, 1))) return rc;  // graph capture outside the timed steps
        if ((rc = sc_batch_step(b, steps))) return rc;
        const double ms = b->last_ms / (double)steps;
        if (entries && q < max_entries) entries[q] = sc_tune_entry{sc::kBuildShapes[q][0], sc::kBuildShapes[q][1], ms};
        if (q == 0 || ms < bms) {
            bms = ms;
            bi = q;
        }
    }
    if (best) *best = bi;
    return sc_batch_set_build_launch(b, sc::kBuildShapes[bi][0], sc::kBuildShapes[bi][1]);
}

}  // extern "C"
