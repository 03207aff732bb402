// sc_api.hip -- C-ABI entry points of libsc (declared in include/sc.h).
//
// A context owns one replica's device state and one HIP stream.  Host
// arrays cross in the reference's AoS layout and original particle order;
// the device keeps SoA state in cell-sorted slot order and maps through the
// per-slot original index.
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sc.h"
#include "sc_internal.h"

namespace sc {
void philox_host(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]);
}

// Device scratch of sc_get_neighbor_lists_noalloc: the reference-layout rows
// and counts are assembled here instead of in buffers allocated per call.
// Every GPU the library runs on holds its own copy for the life of the
// process; concurrent calls on one GPU share it (and race).
#define SC_SCRATCH_WORDS (8u << 20)  // 8-byte words: 64 MiB
#define SC_SCRATCH_FLAG_WORDS 32     // words reserved at the start for flags
static __device__ __align__(256) unsigned long long scratch[SC_SCRATCH_WORDS];

namespace {
// scratch's device address on each GPU (the host-side symbol is a stand-in,
// not an address any kernel may use), looked up once per device
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[256] = {};

int scratch_address(int dev, unsigned long long **out, std::string &err)
{
    if (dev < 0 || dev >= 256) {
        err = "device index out of range for the scratch table";
        return SC_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return SC_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<unsigned long long *>(q);
    }
    *out = g_scratch_ptr[dev];
    return SC_OK;
}
}  // namespace

struct sc_ctx {
    sc_params p{};
    Rep h{};
    Rep *d_rep = nullptr;
    Tile *d_tiles = nullptr;
    int ntiles = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    double *d_noise = nullptr;
    size_t noise_cap = 0;
    double last_ms = 0.0;
    std::vector<void *> allocs;
    bool timing = false;
    std::vector<hipEvent_t> kev;
    sc::RebuildGeom geom{};
    double *d_rold = nullptr;  // reference configuration (original order), sc_set_reference_positions
    double *d_qlm = nullptr;   // Steinhardt scratch: 2 (QL_LMAX+1) doubles and one int per particle
    int *d_nq = nullptr;
    double t_force = 0.0, t_drift = 0.0;
    long long n_force = 0, n_drift = 0;
    RepDyn *h_dyn = nullptr;  // pinned copy of the device state, polled between step chunks
    unsigned long long *scratch = nullptr;  // scratch's address on the context's device
};

namespace {

int fail(sc_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

// A literal message: assigned into the context's reserved text, no allocation.
int fail(sc_ctx *c, int code, const char *msg)
{
    if (c) c->err.assign(msg);
    return code;
}

#define HIPCHK(ctx, expr)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, SC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define KCHK(ctx, expr)                                                            \
    do {                                                                          \
        if ((expr) != 0)                                                          \
            return fail(ctx, SC_ERR_HIP, std::string("kernel launch failed: ") + \
                                             hipGetErrorString(hipGetLastError())); \
    } while (0)

template <typename T>
int dalloc(sc_ctx *c, T **ptr, size_t count)
{
    void *q = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) {
        *ptr = nullptr;
        return fail(c, SC_ERR_OOM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    c->allocs.push_back(q);
    *ptr = static_cast<T *>(q);
    return SC_OK;
}

int push_rep(sc_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->d_rep, &c->h, sizeof(Rep), hipMemcpyHostToDevice, c->s));
    return SC_OK;
}

// Device error bits as a status.
int err_status(sc_ctx *c, int bits, const char *when)
{
    if (bits & SC_EBIT_OVERFLOW)
        return fail(c, SC_ERR_OVERFLOW,
                    std::string(when) + "neighbour list overflow: a particle has more than maxneighbours=" +
                        std::to_string(c->h.cap) + " neighbours within the Verlet radius");
    if (bits & SC_EBIT_NONFINITE) return fail(c, SC_ERR_NONFINITE, std::string(when) + "non-finite particle position");
    if (bits & SC_EBIT_NOISE)
        return fail(c, SC_ERR_STATE, std::string(when) + "a force evaluation had no injected noise block for its step");
    return SC_OK;
}

// Reads the device error bits and turns them into a status.
int check_dyn(sc_ctx *c)
{
    RepDyn d;
    HIPCHK(c, hipMemcpyAsync(&d, c->h.dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return err_status(c, d.err, "");
}

// A context whose device state carries an error from an earlier call is not
// stepped again (its positions or lists are not valid) until
// sc_update_neighbor_lists, which clears the bits, has run -- typically after
// sc_set_state has uploaded a valid state.
int refuse_if_failed(sc_ctx *c, const RepDyn &d)
{
    if (!d.err) return SC_OK;
    return err_status(c, d.err, "an earlier call left this error (rebuild with sc_update_neighbor_lists to continue): ");
}

int read_dyn(sc_ctx *c, RepDyn *d)
{
    HIPCHK(c, hipMemcpyAsync(d, c->h.dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return SC_OK;
}

int upload_field(sc_ctx *c, const double *src, int field)
{
    size_t cnt = (field == FLD_D) ? (size_t)c->h.n : (size_t)c->h.n * c->h.dims;
    HIPCHK(c, hipMemcpyAsync(c->h.stage, src, cnt * sizeof(double), hipMemcpyHostToDevice, c->s));
    KCHK(c, sc::launch_from_host_order(c->d_tiles, c->ntiles, c->d_rep, field, c->s));
    return SC_OK;
}

int download_field(sc_ctx *c, double *dst, int field)
{
    size_t cnt = (field == FLD_D || field == FLD_DISP2) ? (size_t)c->h.n : (size_t)c->h.n * c->h.dims;
    KCHK(c, sc::launch_to_host_order(c->d_tiles, c->ntiles, c->d_rep, field, c->s));
    HIPCHK(c, hipMemcpyAsync(dst, c->h.stage, cnt * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return SC_OK;
}

// Injected noise for the next `blocks` force evaluations, the first of which
// is the device's evaluation number `step0` (RepDyn::step).
int set_noise(sc_ctx *c, const double *noise, int64_t blocks, long long step0)
{
    if (!noise) {
        if (c->h.noise != nullptr) {
            c->h.noise = nullptr;
            return push_rep(c);
        }
        return SC_OK;
    }
    size_t need = (size_t)blocks * c->h.n * c->h.dims;
    if (need > c->noise_cap) {
        if (c->d_noise) {
            HIPCHK(c, hipStreamSynchronize(c->s));
            (void)hipFree(c->d_noise);
            c->d_noise = nullptr;
        }
        void *q = nullptr;
        if (hipMalloc(&q, need * sizeof(double)) != hipSuccess)
            return fail(c, SC_ERR_OOM, "noise buffer allocation failed");
        c->d_noise = static_cast<double *>(q);
        c->noise_cap = need;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_noise, noise, need * sizeof(double), hipMemcpyHostToDevice, c->s));
    c->h.noise = c->d_noise;
    c->h.noise_step0 = step0;
    c->h.noise_blocks = blocks;
    return push_rep(c);
}

int ensure_events(sc_ctx *c, size_t n)
{
    while (c->kev.size() < n) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->kev.push_back(e);
    }
    return SC_OK;
}

// Steps enqueued in chunks ahead of the skin-trigger decision, without the
// five rebuild launches of every step.  Newtonian/Langevin: a step whose
// drift fires the trigger halts the chunk on the device (its force, its
// finish and every later kernel of the chunk return at once, RepDyn::halt),
// the host reads the state after the chunk, runs the rebuild, that step's
// force and finish, and continues.  Brownian (force, finish, drift per step;
// not with kernel timing): the force after a drift that fired the trigger
// halts the chunk; the host runs the rebuild (with its bookkeeping) and
// continues from that step, and also rebuilds when a chunk's last drift
// fired it.  The kernels and their order per step are
// those of the ordinary loop, so the trajectory is bitwise the same.  Chunk
// length: what is left of the last rebuild interval, at least 1, at most 64
// (measured at N=10^6: the early-returning launches a halt leaves in the
// queue cost more than the extra host round trip of a chunk that ends just
// short of the trigger).  Event slots from `base`
// (timing): 0-1 the rebuild step's force, then three per chunk step.
int steps_ahead(sc_ctx *c, int64_t nsteps, bool timed, size_t base)
{
    const int nt = c->ntiles;
    if (!c->h_dyn && hipHostMalloc((void **)&c->h_dyn, sizeof(RepDyn), hipHostMallocDefault) != hipSuccess) {
        c->h_dyn = nullptr;
        return fail(c, SC_ERR_OOM, "hipHostMalloc failed");
    }
    // the state read after each chunk: one copy into pinned memory, one wait
    auto poll = [&](RepDyn *out) -> int {
        HIPCHK(c, hipMemcpyAsync(c->h_dyn, c->h.dyn, sizeof(RepDyn), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        *out = *c->h_dyn;
        return SC_OK;
    };
    RepDyn d;
    int rc = poll(&d);
    if (rc) return rc;
    long long step0 = d.step;
    int64_t left = nsteps, since = 0, interval = 12;
    bool rec_pending = false;
    auto take_rec = [&]() -> int {
        if (!rec_pending) return SC_OK;
        float a = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, c->kev[base], c->kev[base + 1]));
        c->t_force += a;
        c->n_force++;
        rec_pending = false;
        return SC_OK;
    };
    const bool bd = c->h.sys == SC_SYS_BROWNIAN;
    if (bd && timed) return fail(c, SC_ERR_ARG, "Brownian step chunks are not timed");
    while (left > 0) {
        const int64_t m = std::min<int64_t>(std::min<int64_t>(std::max<int64_t>(interval - since, 1), 64), left);
        if (timed && (rc = ensure_events(c, base + 2 + 3 * (size_t)m))) return rc;
        for (int64_t i = 0; bd && i < m; i++) {
            KCHK(c, sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 2, 7, c->s, false, true));
            KCHK(c, sc::launch_force_finish(c->d_rep, 1, 2, 7, c->s, true));
            KCHK(c, sc::launch_drift_brownian(c->d_tiles, nt, c->d_rep, 1, c->s, true));
        }
        for (int64_t i = 0; !bd && i < m; i++) {
            if (timed) HIPCHK(c, hipEventRecord(c->kev[base + 2 + 3 * i], c->s));
            KCHK(c, sc::launch_kick_drift_only(c->d_tiles, nt, c->d_rep, c->s, true));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[base + 3 + 3 * i], c->s));
            KCHK(c, sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 1, 7, c->s, false, true));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[base + 4 + 3 * i], c->s));
            KCHK(c, sc::launch_force_finish(c->d_rep, 1, 1, 7, c->s, true));
        }
        if ((rc = poll(&d))) return rc;
        if ((rc = take_rec())) return rc;
        const int64_t done = (int64_t)(d.step - step0);
        if (done < 0 || done > m || (!d.halt && done != m)) return fail(c, SC_ERR_STATE, "step chunk out of step");
        if (timed) {
            for (int64_t i = 0; i < m && i <= done; i++) {
                if (i == done && !d.halt) break;
                float a = 0.f, b = 0.f;
                HIPCHK(c, hipEventElapsedTime(&b, c->kev[base + 2 + 3 * i], c->kev[base + 3 + 3 * i]));
                c->t_drift += b;
                c->n_drift++;
                if (i < done) {
                    HIPCHK(c, hipEventElapsedTime(&a, c->kev[base + 3 + 3 * i], c->kev[base + 4 + 3 * i]));
                    c->t_force += a;
                    c->n_force++;
                }
            }
        }
        left -= done;
        since += done;
        step0 = d.step;
        if (d.err) break;
        if (bd) {
            // halted at the force after a firing drift, or the last drift fired
            if (d.halt || d.rebuild) {
                KCHK(c, sc::launch_rebuild(c->d_tiles, nt, c->d_rep, 1, c->geom, c->s, true));
                interval = since;
                since = 0;
            }
            continue;
        }
        if (d.halt) {
            // step done+1 drifted and fired the trigger
            KCHK(c, sc::launch_rebuild(c->d_tiles, nt, c->d_rep, 1, c->geom, c->s, false));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[base], c->s));
            KCHK(c, sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 1, 7, c->s, false, false));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[base + 1], c->s));
            KCHK(c, sc::launch_force_finish(c->d_rep, 1, 1, 7, c->s, false));
            rec_pending = timed;
            left -= 1;
            interval = since + 1;
            since = 0;
            step0 += 1;
        }
    }
    if (rec_pending) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        if ((rc = take_rec())) return rc;
    }
    return SC_OK;
}

}  // namespace

extern "C" {

int sc_create(sc_ctx **out, const sc_params *p, const double *D)
{
    if (!out || !p || !D) return SC_ERR_ARG;
    *out = nullptr;
    if (!sc::params_valid(p) || !sc::diameters_valid(D, p->N)) return SC_ERR_ARG;
    sc_ctx *c = new sc_ctx();
    c->p = *p;
    *out = c;
    c->err.reserve(512);  // error text fits without reallocation (the allocation-free call)
    if (hipStreamCreate(&c->s) != hipSuccess) return fail(c, SC_ERR_HIP, "hipStreamCreate failed");
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (int rc0 = scratch_address(dev, &c->scratch, c->err)) return rc0;
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    Rep &h = c->h;
    int rc = sc::replica_init(h, p, D, c->allocs, c->err);
    if (rc != SC_OK) return rc;
    c->ntiles = h.ntiles;
    h.tile0 = 0;
    h.run0 = 0;
    c->geom.max_chunks = (h.ncell + SC_CHUNK - 1) / SC_CHUNK;
    c->geom.nruns = h.build_kind == BUILD_CELLS ? h.nruns : 0;
    c->geom.nwin = h.build_kind == BUILD_WIN ? h.nruns : 0;
    c->geom.any_tile_build = h.build_kind == BUILD_SLOT ? (h.bf_on ? 2 : 1) : 0;
    c->geom.any_rows_build = h.build_kind == BUILD_ROWS ? 1 << h.dims : 0;
    c->geom.flag = reinterpret_cast<const int *>(reinterpret_cast<const char *>(h.dyn) + offsetof(RepDyn, rebuild));
    // default build launch shape by system size (the shapes sc_autotune picked
    // for 3D polydisperse systems of 1000 to 300000 particles): several lanes
    // per slot cut the latency of a small system's rebuild, one lane per slot
    // with the per-slot radius suits a large one; the lists are the same
    sc::default_build_shape(h.n, c->geom.build_bs, c->geom.build_lpp);
    rc = SC_OK;
    rc |= dalloc(c, &c->d_rep, 1);
    rc |= dalloc(c, &c->d_tiles, c->ntiles);
    rc |= dalloc(c, &c->geom.plan, sc::plan_ints(1, c->ntiles, c->geom.nwin));
    if (rc != SC_OK) return SC_ERR_OOM;
    std::vector<Tile> tiles(c->ntiles);
    for (int t = 0; t < c->ntiles; t++) tiles[t] = Tile{0, t * SC_TB, t, 0};
    HIPCHK(c, hipMemcpy(c->d_tiles, tiles.data(), sizeof(Tile) * c->ntiles, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_rep, &h, sizeof(Rep), hipMemcpyHostToDevice));
    return SC_OK;
}

void sc_destroy(sc_ctx *c)
{
    if (!c) return;
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (void *q : c->allocs) (void)hipFree(q);
    if (c->d_noise) (void)hipFree(c->d_noise);
    if (c->h_dyn) (void)hipHostFree(c->h_dyn);
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

const char *sc_last_error(const sc_ctx *c) { return c ? c->err.c_str() : "null context"; }

int sc_set_state(sc_ctx *c, const double *r, const double *v, const double *F, const double *D,
                 const double *disp)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    if (D && !sc::diameters_valid(D, c->h.n)) return fail(c, SC_ERR_ARG, "diameters must be finite and positive");
    int rc;
    if (r && (rc = upload_field(c, r, FLD_R))) return rc;
    if (v && (rc = upload_field(c, v, FLD_V))) return rc;
    if (F && (rc = upload_field(c, F, FLD_F))) return rc;
    if (D) {
        // the windowed build bounds every pair's half-list radius by the
        // largest diameter (swap moves keep the set of diameters), and the
        // mixed form's fixed-point diameters need a scale that holds it
        // (rescaled before the upload, which rewrites every diameter)
        float dm = c->h.dmax, dn = c->h.dmin;
        double dmx = 0.0;
        for (int64_t i = 0; i < c->h.n; i++) {
            dm = std::max(dm, (float)D[i]);
            dn = std::min(dn, (float)D[i]);
            dmx = std::max(dmx, D[i]);
        }
        if (c->h.spos && dmx * c->h.dqinv >= 4294967295.0) {
            sc::set_diameter_scale(c->h, dmx);
            if ((rc = push_rep(c))) return rc;
        }
        if ((rc = upload_field(c, D, FLD_D))) return rc;
        if (dm > c->h.dmax || dn < c->h.dmin) {
            c->h.dmax = dm;
            c->h.dmin = dn;
            // the prefilter's half-list margin follows the diameter range
            // (beyond zeta (max D - min D) = 1/2 the FP64 sweep takes over)
            sc::build_prefilter_setup(c->h);
            if (c->h.build_kind == BUILD_SLOT) c->geom.any_tile_build = c->h.bf_on ? 2 : 1;
            // beyond zeta * max(D) = 2 the windowed build's FP32 margin no
            // longer covers sigma_ij: continue with the exact per-slot sweep
            if (c->h.build_kind == BUILD_WIN && c->h.pot == SC_POT_BERTHIER && c->h.zeta * (double)dm > 2.0) {
                c->h.build_kind = BUILD_SLOT;
                c->geom.nwin = 0;
                c->geom.any_tile_build = 1;
            }
            if ((rc = push_rep(c))) return rc;
        }
    }
    if (disp && (rc = upload_field(c, disp, FLD_DISP))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return SC_OK;
}

int sc_get_state(sc_ctx *c, double *r, double *v, double *F, double *disp, double *disp2)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    int rc;
    if (r && (rc = download_field(c, r, FLD_R))) return rc;
    if (v && (rc = download_field(c, v, FLD_V))) return rc;
    if (F && (rc = download_field(c, F, FLD_F))) return rc;
    if (disp && (rc = download_field(c, disp, FLD_DISP))) return rc;
    if (disp2 && (rc = download_field(c, disp2, FLD_DISP2))) return rc;
    return SC_OK;
}

int sc_update_neighbor_lists(sc_ctx *c)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    KCHK(c, sc::launch_request_rebuild(c->d_rep, 1, c->s));
    KCHK(c, sc::launch_rebuild(c->d_tiles, c->ntiles, c->d_rep, 1, c->geom, c->s));
    return check_dyn(c);
}

int sc_get_neighbor_lists(sc_ctx *c, int64_t maxneigh, int64_t *nfull, int64_t *full, int64_t *nhalf,
                          int64_t *half)
{
    if (!c || !c->d_rep || maxneigh < 1) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    const int n = c->h.n, cap = c->h.cap;
    if (nfull || full) {
        int *rows = nullptr, *cnts = nullptr;
        if (hipMalloc(&rows, sizeof(int) * (size_t)n * cap) != hipSuccess ||
            hipMalloc(&cnts, sizeof(int) * (size_t)n) != hipSuccess) {
            (void)hipFree(rows);
            return fail(c, SC_ERR_OOM, "list export buffers");
        }
        std::vector<int> hr((size_t)n * cap), hc(n);
        KCHK(c, sc::launch_full_rows(c->d_tiles, c->ntiles, c->d_rep, rows, cnts, cap, c->s));
        HIPCHK(c, hipMemcpyAsync(hr.data(), rows, sizeof(int) * hr.size(), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipMemcpyAsync(hc.data(), cnts, sizeof(int) * n, hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        (void)hipFree(rows);
        (void)hipFree(cnts);
        for (int i = 0; i < n; i++) {
            if (hc[i] > cap || hc[i] > maxneigh)
                return fail(c, SC_ERR_OVERFLOW, "full list row longer than maxneigh");
            if (nfull) nfull[i] = hc[i];
            if (full)
                for (int k = 0; k < hc[i]; k++) full[(size_t)i * maxneigh + k] = (int64_t)hr[(size_t)i * cap + k] + 1;
        }
    }
    if (nhalf || half) {
        std::vector<int> nn(n), orig(n), nb((size_t)n * cap);
        HIPCHK(c, hipMemcpyAsync(nn.data(), c->h.nn, sizeof(int) * n, hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipMemcpyAsync(orig.data(), c->h.orig, sizeof(int) * n, hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipMemcpyAsync(nb.data(), c->h.nbr, sizeof(int) * nb.size(), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        std::vector<int64_t> cnt(n, 0);
        for (int k = 0; k < n; k++) {
            const int a = orig[k];
            for (int q = 0; q < nn[k]; q++) {
                const int b = orig[nb[(size_t)q * n + k]];
                if (a < b) {
                    if (cnt[a] >= maxneigh) return fail(c, SC_ERR_OVERFLOW, "half list row longer than maxneigh");
                    if (half) half[(size_t)a * maxneigh + cnt[a]] = (int64_t)b + 1;
                    cnt[a]++;
                }
            }
        }
        if (nhalf)
            for (int i = 0; i < n; i++) nhalf[i] = cnt[i];
    }
    return SC_OK;
}

int sc_get_neighbor_lists_noalloc(sc_ctx *c, int64_t maxneigh, int64_t *nfull, int64_t *full, int64_t *nhalf,
                                  int64_t *half)
{
    if (!c || !c->d_rep || maxneigh < 1 || maxneigh > 1000000) return SC_ERR_ARG;
    const int64_t n = c->h.n;
    const bool want_full = nfull || full, want_half = nhalf || half;
    const uint64_t need = SC_SCRATCH_FLAG_WORDS + (uint64_t)(want_full + want_half) * (uint64_t)n * (maxneigh + 1);
    if (need > SC_SCRATCH_WORDS)
        return fail(c, SC_ERR_ARG, "N * (maxneigh + 1) too large for the 64 MiB export scratch "
                                   "(at most 8388576 words per list requested); use sc_get_neighbor_lists");
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    int *flag = reinterpret_cast<int *>(c->scratch);
    long long *base = reinterpret_cast<long long *>(c->scratch + SC_SCRATCH_FLAG_WORDS);
    int over[2] = {0, 0};
    for (int h = 0; h < 2; h++) {
        if (!(h ? want_half : want_full)) continue;
        long long *rows = base, *cnts = base + (size_t)n * maxneigh;
        base += (size_t)n * (maxneigh + 1);
        KCHK(c, sc::launch_ref_rows(c->d_tiles, c->ntiles, c->d_rep, h, rows, cnts, (int)maxneigh, flag + h, c->s));
        int64_t *dst_rows = h ? half : full, *dst_cnt = h ? nhalf : nfull;
        if (dst_rows)
            HIPCHK(c, hipMemcpyAsync(dst_rows, rows, sizeof(int64_t) * (size_t)n * maxneigh, hipMemcpyDeviceToHost, c->s));
        if (dst_cnt) HIPCHK(c, hipMemcpyAsync(dst_cnt, cnts, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost, c->s));
    }
    HIPCHK(c, hipMemcpyAsync(over, flag, sizeof(over), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    // each requested list cleared its own flag word first; the other word is stale
    if (want_full && over[0]) return fail(c, SC_ERR_OVERFLOW, "full list row longer than maxneigh");
    if (want_half && over[1]) return fail(c, SC_ERR_OVERFLOW, "half list row longer than maxneigh");
    return SC_OK;
}

int sc_calculate_new_forces(sc_ctx *c, const double *noise)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    if ((rc = refuse_if_failed(c, d))) return rc;
    if ((rc = set_noise(c, c->h.sys != SC_SYS_NEWTONIAN ? noise : nullptr, 1, d.step))) return rc;
    KCHK(c, sc::launch_force(c->d_tiles, c->ntiles, c->d_rep, 1, 0, 7, c->s));
    return check_dyn(c);
}

int sc_do_time_steps(sc_ctx *c, int64_t nsteps, const double *noise)
{
    if (!c || !c->d_rep || nsteps < 0) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    if ((rc = refuse_if_failed(c, d))) return rc;
    if ((rc = set_noise(c, c->h.sys != SC_SYS_NEWTONIAN ? noise : nullptr, nsteps, d.step))) return rc;
    const int nt = c->ntiles;
    const bool timed = c->timing;
    if (timed && (rc = ensure_events(c, 4 * (size_t)nsteps))) return rc;
    size_t ev = 0;
    HIPCHK(c, hipEventRecord(c->ev0, c->s));
    // Newtonian/Langevin without rescaling: chunks enqueued ahead of the
    // rebuild decision (steps_ahead); a rebuild requested before this call
    // goes through one ordinary step first
    const bool ahead = c->h.sys == SC_SYS_BROWNIAN ? !timed : !c->h.rescale;
    const int64_t plain = ahead ? ((d.rebuild || d.halt) ? std::min<int64_t>(1, nsteps) : 0) : nsteps;
    for (int64_t s = 0; s < plain; s++) {
        if (c->h.sys == SC_SYS_BROWNIAN) {
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            KCHK(c, sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 2, 7, c->s));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            KCHK(c, sc::launch_drift_brownian(c->d_tiles, nt, c->d_rep, 1, c->s));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            // bookkeeping by the next step's force finish (one launch less per step)
            KCHK(c, sc::launch_rebuild(c->d_tiles, nt, c->d_rep, 1, c->geom, c->s, false));
        } else {
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            KCHK(c, sc::launch_kick_drift_only(c->d_tiles, nt, c->d_rep, c->s));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            KCHK(c, sc::launch_rebuild(c->d_tiles, nt, c->d_rep, 1, c->geom, c->s, false));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            KCHK(c, sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 1, 7, c->s, false));
            if (timed) HIPCHK(c, hipEventRecord(c->kev[ev++], c->s));
            KCHK(c, sc::launch_force_finish(c->d_rep, 1, 1, 7, c->s));
            if (c->h.rescale) KCHK(c, sc::launch_ekin_com(c->d_tiles, nt, c->d_rep, 1, c->s));
        }
    }
    if (ahead && nsteps > plain) {
        // a Brownian step ends with its rebuild, whose bookkeeping the next
        // force finish would do: do it now, the chunks expect a clear flag
        if (plain > 0 && c->h.sys == SC_SYS_BROWNIAN) KCHK(c, sc::launch_rebuild_done(c->d_rep, 1, c->s));
        if ((rc = steps_ahead(c, nsteps - plain, timed, ev))) return rc;
    }
    // the last Brownian step's rebuild bookkeeping (no next force finish)
    if (c->h.sys == SC_SYS_BROWNIAN && nsteps > 0) KCHK(c, sc::launch_rebuild_done(c->d_rep, 1, c->s));
    KCHK(c, sc::launch_finalize(c->d_tiles, nt, c->d_rep, 1, c->s));
    HIPCHK(c, hipEventRecord(c->ev1, c->s));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (timed) {
        // Newtonian/Langevin: [drift][rebuild][force]; Brownian: [force][drift][rebuild]
        for (size_t q = 0; q < ev; q += 3) {
            float a = 0.f, b = 0.f;
            if (c->h.sys == SC_SYS_BROWNIAN) {
                HIPCHK(c, hipEventElapsedTime(&a, c->kev[q], c->kev[q + 1]));
                HIPCHK(c, hipEventElapsedTime(&b, c->kev[q + 1], c->kev[q + 2]));
                c->t_force += a;
                c->t_drift += b;
            } else {
                HIPCHK(c, hipEventElapsedTime(&b, c->kev[q], c->kev[q + 1]));
                HIPCHK(c, hipEventElapsedTime(&a, c->kev[q + 2], c->kev[q + 3]));
                c->t_force += a;
                c->t_drift += b;
                q += 1;
            }
            c->n_force++;
            c->n_drift++;
        }
    }
    return check_dyn(c);
}

int sc_full_energy(sc_ctx *c, double *out)
{
    if (!c || !c->d_rep || !out) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    KCHK(c, sc::launch_energy(c->d_tiles, c->ntiles, c->d_rep, 1, c->s));
    if ((rc = read_dyn(c, &d))) return rc;
    *out = d.result[0];
    return SC_OK;
}

int sc_potential_energies(sc_ctx *c, double *E)
{
    if (!c || !c->d_rep || !E) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    KCHK(c, sc::launch_energy(c->d_tiles, c->ntiles, c->d_rep, 1, c->s, c->h.stage));
    HIPCHK(c, hipMemcpyAsync(E, c->h.stage, sizeof(double) * c->h.n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return SC_OK;
}

int sc_energy_no_neigh(sc_ctx *c, double *out)
{
    if (!c || !c->d_rep || !out) return SC_ERR_ARG;
    KCHK(c, sc::launch_energy_nn(c->d_tiles, c->ntiles, c->d_rep, 1, c->s));
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    *out = d.result[0];
    return SC_OK;
}

int sc_set_reference_positions(sc_ctx *c)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    int rc;
    if (!c->d_rold && (rc = dalloc(c, &c->d_rold, (size_t)3 * c->h.n))) return rc;
    KCHK(c, sc::launch_set_rold(c->d_tiles, c->ntiles, c->d_rep, c->d_rold, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return SC_OK;
}

int sc_f2_self(sc_ctx *c, double kx, double *out)
{
    if (!c || !c->d_rep || !out) return SC_ERR_ARG;
    if (!c->d_rold) return fail(c, SC_ERR_STATE, "no reference configuration (sc_set_reference_positions)");
    KCHK(c, sc::launch_f2s(c->d_tiles, c->ntiles, c->d_rep, 1, c->d_rold, kx, c->s));
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    *out = d.result[0];
    return SC_OK;
}

int sc_steinhardt(sc_ctx *c, int32_t l, double q_cutoff, int32_t averaged, double *out)
{
    if (!c || !c->d_rep || !out) return SC_ERR_ARG;
    if (c->h.dims != 3) return fail(c, SC_ERR_ARG, "Steinhardt order parameters are defined for dims == 3");
    if (l < 1 || l > sc::QL_LMAX) return fail(c, SC_ERR_ARG, "Steinhardt order l must be 1.." + std::to_string(sc::QL_LMAX));
    if (!(q_cutoff >= 0.0)) return fail(c, SC_ERR_ARG, "q_cutoff must be non-negative");
    // the reference's neighbours come from the full list, i.e. within R_v
    const double qc2 = std::min(q_cutoff * q_cutoff, c->h.Rv2);
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    if (!c->d_qlm) {
        if ((rc = dalloc(c, &c->d_qlm, (size_t)2 * (sc::QL_LMAX + 1) * c->h.n))) return rc;
        if ((rc = dalloc(c, &c->d_nq, (size_t)c->h.n))) return rc;
    }
    KCHK(c, sc::launch_steinhardt(c->d_tiles, c->ntiles, c->d_rep, 1, l, qc2, averaged ? 1 : 0,
                                  c->d_qlm, c->d_nq, c->s));
    if ((rc = read_dyn(c, &d))) return rc;
    *out = d.result[0];
    return SC_OK;
}

int sc_kinetic_energy(sc_ctx *c, double *out)
{
    if (!c || !c->d_rep || !out) return SC_ERR_ARG;
    KCHK(c, sc::launch_ekin(c->d_tiles, c->ntiles, c->d_rep, 1, 0, c->s));
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    *out = d.result[1];
    return SC_OK;
}

int sc_rescale_velocities(sc_ctx *c)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    KCHK(c, sc::launch_ekin(c->d_tiles, c->ntiles, c->d_rep, 1, 1, c->s));
    KCHK(c, sc::launch_finalize(c->d_tiles, c->ntiles, c->d_rep, 1, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return SC_OK;
}

int sc_rebuild_count(sc_ctx *c, int64_t *out, int64_t set)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (out) *out = d.rebuilds;
    if (set >= 0) {
        long long v = set;
        HIPCHK(c, hipMemcpyAsync(reinterpret_cast<char *>(c->h.dyn) + offsetof(RepDyn, rebuilds), &v, sizeof(v),
                                 hipMemcpyHostToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    return SC_OK;
}

int sc_minimize(sc_ctx *c, int64_t iters, double dt, double max_step, double *maxforce)
{
    if (!c || !c->d_rep || iters < 0 || !(dt > 0.0) || !(max_step > 0.0)) return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    if ((rc = refuse_if_failed(c, d))) return rc;
    const int nt = c->ntiles;
    // pair forces only: a Newtonian view of the replica for the force kernel
    Rep saved = c->h;
    c->h.sys = SC_SYS_NEWTONIAN;
    if ((rc = push_rep(c))) return rc;
    for (int64_t it = 0; it < iters; it++) {
        KCHK(c, sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 0, 7, c->s));
        KCHK(c, sc::launch_min_scale(c->d_tiles, nt, c->d_rep, 1, dt, max_step, c->s));
        KCHK(c, sc::launch_min_move(c->d_tiles, nt, c->d_rep, 1, dt, c->s));
        KCHK(c, sc::launch_rebuild(c->d_tiles, nt, c->d_rep, 1, c->geom, c->s));
    }
    c->h = saved;
    if ((rc = push_rep(c))) return rc;
    if ((rc = read_dyn(c, &d))) return rc;
    if (maxforce) *maxforce = d.result[3];
    return check_dyn(c);
}

// One minimiser iteration: force, step scale, move, skin-triggered rebuild.
static int min_iteration(sc_ctx *c, double dt, double max_step)
{
    const int nt = c->ntiles;
    if (sc::launch_force(c->d_tiles, nt, c->d_rep, 1, 0, 7, c->s) ||
        sc::launch_min_scale(c->d_tiles, nt, c->d_rep, 1, dt, max_step, c->s) ||
        sc::launch_min_move(c->d_tiles, nt, c->d_rep, 1, dt, c->s) ||
        sc::launch_rebuild(c->d_tiles, nt, c->d_rep, 1, c->geom, c->s))
        return fail(c, SC_ERR_HIP, std::string("kernel launch failed: ") + hipGetErrorString(hipGetLastError()));
    return SC_OK;
}

int sc_minimize_until(sc_ctx *c, int64_t max_iters, double dt, double max_step, double tolerance,
                      int64_t *iterations, double *maxforce, int32_t *converged)
{
    if (!c || !c->d_rep || max_iters < 0 || !(dt > 0.0) || !(max_step > 0.0) || !(tolerance >= 0.0))
        return SC_ERR_ARG;
    RepDyn d;
    int rc = read_dyn(c, &d);
    if (rc) return rc;
    if (!d.have_lists) return fail(c, SC_ERR_STATE, "neighbour lists have not been built");
    if ((rc = refuse_if_failed(c, d))) return rc;
    // Each iteration checks max(max|F|, step) <= tolerance on the device after
    // its move (FIN_MINSCALE); from then on the force, scale and move kernels
    // of the remaining launched iterations return at once.  The host launches
    // growing chunks and polls the stop state between them, so at most one
    // chunk of no-op iterations is enqueued past convergence.
    Rep saved = c->h;
    c->h.sys = SC_SYS_NEWTONIAN;
    c->h.min_tol = tolerance;
    rc = push_rep(c);
    if (!rc && sc::launch_min_mark(c->d_rep, 1, true, c->s))
        rc = fail(c, SC_ERR_HIP, "kernel launch failed");
    int64_t done = 0, chunk = 16;
    bool conv = false;
    while (!rc && done < max_iters) {
        const int64_t k = std::min(chunk, max_iters - done);
        for (int64_t it = 0; !rc && it < k; it++) rc = min_iteration(c, dt, max_step);
        done += k;
        if (!rc) rc = read_dyn(c, &d);
        if (!rc) conv = d.min_state != 0;
        if (rc || conv || d.err) break;
        chunk = std::min<int64_t>(chunk * 2, 1024);
    }
    if (!rc && sc::launch_min_mark(c->d_rep, 1, false, c->s)) rc = fail(c, SC_ERR_HIP, "kernel launch failed");
    c->h = saved;
    int rc2 = push_rep(c);
    if (rc) return rc;
    if (rc2) return rc2;
    if ((rc = read_dyn(c, &d))) return rc;
    if (iterations) *iterations = d.min_iters;
    if (maxforce) *maxforce = d.result[3];
    if (converged) *converged = conv ? 1 : 0;
    return check_dyn(c);
}

int sc_copy_bandwidth(int64_t bytes, int32_t reps, double *gbps)
{
    if (bytes < 1024 || reps < 1 || !gbps) return SC_ERR_ARG;
    const size_t n = (size_t)bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SC_OK;
    if (hipMalloc(&a, n * sizeof(double4)) != hipSuccess || hipMalloc(&b, n * sizeof(double4)) != hipSuccess)
        rc = SC_ERR_OOM;
    if (rc == SC_OK && (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, n * sizeof(double4), s) != hipSuccess))
        rc = SC_ERR_HIP;
    if (rc == SC_OK) rc = sc::launch_copy(a, b, n, s) ? SC_ERR_HIP : SC_OK;  // warm-up
    if (rc == SC_OK && hipEventRecord(e0, s) != hipSuccess) rc = SC_ERR_HIP;
    for (int r = 0; rc == SC_OK && r < reps; r++) rc = sc::launch_copy(a, b, n, s) ? SC_ERR_HIP : SC_OK;
    float ms = 0.0f;
    if (rc == SC_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.0f)))
        rc = SC_ERR_HIP;
    if (rc == SC_OK) *gbps = 2.0 * (double)(n * sizeof(double4)) * reps / ((double)ms * 1e-3) / 1e9;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int sc_last_elapsed_ms(sc_ctx *c, double *ms)
{
    if (!c || !ms) return SC_ERR_ARG;
    *ms = c->last_ms;
    return SC_OK;
}

int sc_set_kernel_timing(sc_ctx *c, int on)
{
    if (!c) return SC_ERR_ARG;
    c->timing = on != 0;
    c->t_force = c->t_drift = 0.0;
    c->n_force = c->n_drift = 0;
    return SC_OK;
}

int sc_kernel_times(sc_ctx *c, double *force_ms, int64_t *force_launches, double *drift_ms, int64_t *drift_launches)
{
    if (!c) return SC_ERR_ARG;
    if (force_ms) *force_ms = c->t_force;
    if (force_launches) *force_launches = c->n_force;
    if (drift_ms) *drift_ms = c->t_drift;
    if (drift_launches) *drift_launches = c->n_drift;
    return SC_OK;
}

int sc_neighbor_stats(sc_ctx *c, int64_t *total, int64_t *max_row)
{
    if (!c || !c->d_rep) return SC_ERR_ARG;
    std::vector<int> nn(c->h.n);
    HIPCHK(c, hipMemcpyAsync(nn.data(), c->h.nn, sizeof(int) * nn.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    int64_t t = 0, m = 0;
    for (int x : nn) {
        t += x;
        m = std::max<int64_t>(m, x);
    }
    if (total) *total = t;
    if (max_row) *max_row = m;
    return SC_OK;
}

void sc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { sc::philox_host(ctr, key, out); }

int sc_set_build_launch(sc_ctx *c, int32_t block_size, int32_t lanes_per_particle)
{
    if (!c || !sc::build_shape_valid(block_size, lanes_per_particle)) return SC_ERR_ARG;
    c->geom.build_bs = block_size;
    c->geom.build_lpp = lanes_per_particle;
    return SC_OK;
}

int sc_autotune(sc_ctx *c, int64_t steps, sc_tune_entry *entries, int32_t max_entries, int32_t *best)
{
    if (!c || !c->d_rep || steps < 1 || max_entries < 0) return SC_ERR_ARG;
    int bi = 0, rc;
    double bms = 0.0;
    for (int q = 0; q < SC_TUNE_SHAPES; q++) {
        if ((rc = sc_set_build_launch(c, sc::kBuildShapes[q][0], sc::kBuildShapes[q][1]))) return rc;
        if ((rc = sc_do_time_steps(c, 1, nullptr))) return rc;
        if ((rc = sc_do_time_steps(c, steps, nullptr))) return rc;
        const double ms = c->last_ms / (double)steps;
        if (entries && q < max_entries) entries[q] = sc_tune_entry{sc::kBuildShapes[q][0], sc::kBuildShapes[q][1], ms};
        if (q == 0 || ms < bms) {
            bms = ms;
            bi = q;
        }
    }
    if (best) *best = bi;
    return sc_set_build_launch(c, sc::kBuildShapes[bi][0], sc::kBuildShapes[bi][1]);
}

}  // extern "C"
