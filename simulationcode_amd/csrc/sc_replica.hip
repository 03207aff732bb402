// sc_replica.hip -- device state of one replica (one periodic system), shared
// by the single-system context (sc_create) and the replica batch.
//
// initialize_neighbor_struct (src/DataStructs.jl:349-378) fixes the Verlet
// radius R_v = skin + force_cutoff * max(D); the cell grid, the tile count and
// every device buffer follow from it and from N.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sc.h"
#include "sc_internal.h"

namespace sc {

bool params_valid(const sc_params *p)
{
    return !((p->dims != 2 && p->dims != 3) || (p->potential != SC_POT_LJ && p->potential != SC_POT_BERTHIER) ||
             p->system < 0 || p->system > 2 || (p->precision != SC_PREC_FP64 && p->precision != SC_PREC_MIXED) ||
             p->N < 2 || p->N > (int64_t)1 << 30 || p->maxneighbours < 1 || p->maxneighbours > 100000 ||
             !(p->box_size > 0.0) || !(p->skin_distance >= 0.0) || !(p->force_cutoff > 0.0) ||
             p->cell_divisor < 0 || p->cell_divisor > 4 || p->build_mode < 0 || p->build_mode > 4);
}

// Diameters set the Verlet radius, the cell grid and every pair's sigma_ij:
// each must be a finite positive number (checked before any device work).
bool diameters_valid(const double *D, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        if (!(D[i] > 0.0) || !std::isfinite(D[i])) return false;
    return true;
}

// Cell grid of a replica: cells of side >= R_v / k with a (2k+1)^dims
// stencil; a hair of slack so floating-point cell assignment can never drop
// a pair at exactly R_v.  Fewer than 2k+1 cells per side => one cell.
static void cell_grid(const sc_params *p, double Rv, long long &nc, int &srange)
{
    const int k = p->cell_divisor > 0 ? p->cell_divisor : 2;
    nc = (long long)std::floor(k * p->box_size / (Rv * (1.0 + 1e-10)));
    srange = k;
    if (nc < 2 * k + 1) {
        nc = 1;
        srange = 0;
    }
}

int64_t replica_bytes(const sc_params *p, const double *D)
{
    double maxD = D[0];
    for (int64_t i = 1; i < p->N; i++) maxD = std::max(maxD, D[i]);
    const double Rv = p->skin_distance + p->force_cutoff * maxD;
    long long nc;
    int srange;
    cell_grid(p, Rv, nc, srange);
    const long long ncell = nc * nc * (p->dims == 3 ? nc : 1);
    const long long n = p->N, d = p->dims;
    const long long ntiles = (n + SC_TB - 1) / SC_TB;
    long long b = 0;
    b += 2 * n * 32;                  // pos, pos_alt
    b += 4 * n * d * 8;               // v, v_alt, F, disp
    b += 5 * n * 4;                   // orig, orig_alt, cell_of, rank, key
    b += (2 * ncell + 1) * 4 + 4 * (1 + SC_CHUNK_SPLIT) * ((ncell + SC_CHUNK - 1) / SC_CHUNK);
    b += n * 4 + n * p->maxneighbours * 4;  // nn, nbr
    if (p->precision == SC_PREC_MIXED) b += n * 16;  // spos
    if (p->build_mode == 0 && srange >= 1) b += n * 16;  // bpos
    b += ntiles * 8 * 8 + n * 3 * 8 + 4096;  // partials, staging, dyn
    return b;
}

// Fixed-point diameters of the mixed form: w = D * 2^(31-e) with D_max <
// 2^e, so w < 2^31 and diameters up to twice the largest still fit 32 bits;
// D = w * 2^(e-31) then resolves D to 2^-30 of its power-of-two bound
// (about 1e-9 relative at D ~ 1).
void set_diameter_scale(Rep &h, double maxD)
{
    int e = 0;
    (void)std::frexp(maxD > 0.0 ? maxD : 1.0, &e);  // maxD = f 2^e, f in [0.5, 1)
    h.dqinv = std::ldexp(1.0, 31 - e);
    h.dq = std::ldexp(1.0, e - 31);
}

unsigned diameter_fixed(double D, double dqinv) { return (unsigned)(long long)std::rint(D * dqinv); }

// Margins of the per-slot build's fixed-point prefilter.  A coordinate's
// u = rint(x S) (S = 2^32 / L, x S rounded in FP64) lies within 0.51 of x S,
// so the wrapping difference of two lies within 1.02 of the minimum-image
// separation times S; its FP32 value adds 2^-24 relative, and the squares and
// the sum in FP32 three roundings of 2^-24.  For a separation d, D2 = d^2 S^2:
//   |d2_32 - D2| <= 6 2^-24 D2 + 2.1 sqrt(dims) sqrt(D2) + 1.1 dims =: M(D2),
// increasing in D2, and the FP64 d2 of the same positions differs from D2 by
// far less.  Four times M(R_v^2 S^2) is taken on both sides of R_v^2 S^2.
// The half-list radius (rcut sigma_ij + skin) S is evaluated in FP32 from
// FP32 diameters as fma(c1, (D_i + D_j) t, c2), t = fma(-zeta, |D_i - D_j|, 1),
// with the constants c1 = rcut S / 2 and c2 = skin S rounded once each, and
// then squared.  In units of u = 2^-24, while t >= 1/2: |D_i - D_j| is off
// by at most 2 u max D (diameter roundings and the subtraction), so t by
// (6 zeta max D + 1) u relative (zeta's rounding, the fused rounding, and
// the factor 2 from t >= 1/2); D_i + D_j by 2 u, their product by
// (4 + 6 zeta max D) u, the radius by (6 + 6 zeta max D) u (c1's rounding
// and the fused rounding; the skin term's error is smaller) and its square
// by (13 + 12 zeta max D) u.  The margin keeps eight times the larger
// (23 + 16 zeta max D) u, plus the absolute 4 M(R_v^2 S^2).  The LJ radius is one
// FP64 constant, given 2^-20 relative slack and the same absolute margin.
void build_prefilter_setup(Rep &h)
{
    const double S = h.uq, Rv = std::sqrt(h.Rv2), u = std::ldexp(1.0, -24);
    const double RvS2 = h.Rv2 * S * S;
    const double m4 = 4.0 * (6.0 * u * RvS2 + 2.1 * std::sqrt((double)h.dims) * Rv * S + 1.1 * h.dims);
    h.bf_lo = (float)(RvS2 - m4);
    h.bf_hi = (float)(RvS2 + m4);
    h.bf_m = (float)m4;
    h.bf_c1 = (float)(0.5 * h.rcut * S);
    h.bf_c2 = (float)(h.skin * S);
    h.bf_zeta = (float)h.zeta;
    const double dmax = h.dmax, dmin = h.dmin;
    const double eh = 8.0 * (23.0 + 16.0 * std::fabs(h.zeta) * dmax) * u;
    h.bf_hlo = (float)(1.0 - eh);
    h.bf_hhi = (float)(1.0 + eh);
    const double rh = h.rcut * h.lj_md + h.skin, rh2 = rh * rh * S * S;
    h.bf_rh2lo = (float)(rh2 * (1.0 - std::ldexp(1.0, -20)) - m4);
    h.bf_rh2hi = (float)(rh2 * (1.0 + std::ldexp(1.0, -20)) + m4);
    bool ok = h.bpos != nullptr && h.srange >= 1 && h.build_kind == BUILD_SLOT;
    if (h.pot == SC_POT_BERTHIER) ok = ok && h.zeta >= 0.0 && h.zeta * (dmax - dmin) <= 0.5;
    h.bf_on = ok ? 1 : 0;
}

static int alloc_into(std::vector<void *> &allocs, void **ptr, size_t bytes, std::string &err)
{
    void *q = nullptr;
    if (bytes == 0) bytes = 1;
    if (hipMalloc(&q, bytes) != hipSuccess) {
        *ptr = nullptr;
        err = "device allocation of " + std::to_string(bytes) + " bytes failed";
        return SC_ERR_OOM;
    }
    allocs.push_back(q);
    *ptr = q;
    return SC_OK;
}

template <typename T>
static int A(std::vector<void *> &allocs, T **ptr, size_t count, std::string &err)
{
    return alloc_into(allocs, reinterpret_cast<void **>(ptr), count * sizeof(T), err);
}

// Windowed build runs: cx x cy x cz own cells (cz = 1 in 2D) whose window
// (own cells plus srange cells on every side) holds about 1200 particles on
// average, well inside the SC_WIN_MAX LDS slots (a denser window falls back
// to global-memory candidates).  Among the shapes that fit, the one with the
// most own cells per window cell (fewest staged loads per particle) wins.
//
// FP32 prefilter margins: window-relative coordinates are below
// xmax = (largest window side + 1) cells, so each rounded coordinate is off
// by <= xmax 2^-24 and a separation component by e1 <= 2 xmax 2^-24 +
// R_v 2^-24; near R_v the FP32 d^2 is then within 3 (2 R_v e1 + e1^2) +
// 4 R_v^2 2^-24 of the FP64 d^2.  Four times that bound is used, and the
// half-list radius (FP32 sigma_ij) gets a further 4 R_v^2 2^-20.
static void window_geometry(Rep &h, int dims, double Rv)
{
    const int nc = h.nc, s = h.srange;
    const double per_cell = (double)h.n / (double)h.ncell;
    const int lim = std::min(nc, 32);
    double best = -1.0;
    int bx = 1, by = 1, bz = 1;
    for (int cx = 1; cx <= lim; cx++)
        for (int cy = 1; cy <= lim; cy++)
            for (int cz = 1; cz <= (dims == 3 ? lim : 1); cz++) {
                const long long wc = (long long)(cx + 2 * s) * (cy + 2 * s) * (dims == 3 ? cz + 2 * s : 1);
                if (wc > SC_WIN_CELLS || (long long)cy * cz > SC_WIN_ROWS) continue;
                if (wc * per_cell > 1200.0 && !(cx == 1 && cy == 1 && cz == 1)) continue;
                const double own = (double)cx * cy * cz;
                const double score = own / (double)wc + 1e-6 * own;
                if (score > best) {
                    best = score;
                    bx = cx;
                    by = cy;
                    bz = cz;
                }
            }
    h.cx = bx;
    h.cy = by;
    h.cz = bz;
    const long long rx = (nc + bx - 1) / bx, ry = (nc + by - 1) / by, rz = dims == 3 ? (nc + bz - 1) / bz : 1;
    h.nruns = (int)(rx * ry * rz);
    const int wmax = std::max(bx, std::max(by, bz)) + 2 * s + 1;
    const double xmax = wmax * h.wcell;
    const double u = std::ldexp(1.0, -24);
    const double e1 = 2.0 * xmax * u + Rv * u + 1e-12;
    const double Rv2 = Rv * Rv;
    const double m = 4.0 * (3.0 * (2.0 * Rv * e1 + e1 * e1) + 4.0 * Rv2 * u) + 1e-12;
    h.wlo = (float)(Rv2 - m);
    h.whi = (float)(Rv2 + m);
    h.hmarg = (float)(m + 4.0 * Rv2 * std::ldexp(1.0, -20));
}

int replica_init(Rep &h, const sc_params *p, const double *D, std::vector<void *> &allocs, std::string &err)
{
    std::memset(&h, 0, sizeof(Rep));
    const int n = (int)p->N;
    double maxD = D[0], minD = D[0];
    for (int i = 1; i < n; i++) {
        maxD = std::max(maxD, D[i]);
        minD = std::min(minD, D[i]);
    }
    // initialize_neighbor_struct, src/DataStructs.jl:351-352
    const double Rv = p->skin_distance + p->force_cutoff * maxD;
    if (!(p->box_size > 2.0 * Rv)) {
        err = "box_size must exceed twice the Verlet radius";
        return SC_ERR_ARG;
    }
    h.dims = p->dims;
    h.pot = p->potential;
    h.sys = p->system;
    h.prec = p->precision;
    h.n = n;
    h.cap = (int)p->maxneighbours;
    h.rescale = (p->system == SC_SYS_NEWTONIAN && p->rescale_temperature) ? 1 : 0;
    h.L = p->box_size;
    h.invL = 1.0 / p->box_size;
    h.fc2 = p->force_cutoff2;
    h.rcut = std::sqrt(p->force_cutoff2);  // src/NeighborLists.jl:44
    h.skin = p->skin_distance;
    h.Rv2 = Rv * Rv;
    h.thr = (p->skin_distance / 2.0) * (p->skin_distance / 2.0);
    std::memcpy(&h.thr_bits, &h.thr, sizeof(double));
    h.eps = p->lj_eps;
    h.lj_s2 = p->lj_sigma * p->lj_sigma;
    h.lj_md = p->lj_sigma * p->lj_sigma;  // find_mean_D for LJ returns sigma^2
    h.Uc = p->lj_Uc;
    h.c0 = p->b_c0;
    h.c2 = p->b_c2;
    h.c4 = p->b_c4;
    h.zeta = p->b_zeta;
    h.kBT = p->kBT;
    h.gamma = p->gamma;
    h.dt = p->dt;
    h.m = p->m;
    h.min_tol = -1.0;
    h.seed = p->seed;
    h.uq = std::ldexp(1.0, 32) / p->box_size;
    h.lq = p->box_size * std::ldexp(1.0, -32);
    set_diameter_scale(h, maxD);
    long long nc;
    int srange;
    cell_grid(p, Rv, nc, srange);
    const long long ncell = nc * nc * (p->dims == 3 ? nc : 1);
    if (ncell > (1LL << 27)) {
        err = "cell grid too large";
        return SC_ERR_ARG;
    }
    h.nc = (int)nc;
    h.ncell = (int)ncell;
    h.srange = srange;
    h.cinv = (double)nc / p->box_size;
    h.ntiles = (n + SC_TB - 1) / SC_TB;
    // build kind: the windowed LDS build (build_mode 3, stencils up to +-3),
    // the cell-run LDS build (build_mode 1, stencil +-1), the staged-row LDS
    // build (build_mode 4, any stencil), else (automatic and build_mode 2) the
    // per-slot sweep
    h.wcell = p->box_size / (double)nc;
    h.dmax = (float)maxD;
    h.dmin = (float)minD;
    // the windowed build's FP32 half-list radius stays within its margin while
    // 0 <= zeta * max(D) <= 2 (sigma_ij error <= dmax (6 zeta dmax + 4) 2^-24,
    // threshold error <= R_v^2 (12 zeta dmax + 15) 2^-24 < 4 R_v^2 2^-20);
    // outside that range mode 3 falls back to the exact per-slot sweep
    const bool win_ok = p->potential != SC_POT_BERTHIER || (p->b_zeta >= 0.0 && p->b_zeta * maxD <= 2.0);
    if (srange >= 1 && srange <= 3 && p->build_mode == 3 && win_ok) {
        h.build_kind = BUILD_WIN;
        window_geometry(h, p->dims, Rv);
    } else if (srange == 1 && p->build_mode == 1) {
        h.build_kind = BUILD_CELLS;
        const double per_cell = (double)n / (double)ncell;
        const int cxmax = (p->dims == 3 ? 5 : 19);
        int cx = (int)std::lround(64.0 / std::max(per_cell, 1e-9));
        cx = std::max(1, std::min(cx, std::min(cxmax, (int)nc - 2)));
        h.cx = cx;
        const long long rpr = (nc + cx - 1) / cx;
        h.nruns = (int)(rpr * nc * (p->dims == 3 ? nc : 1));
    } else if (srange >= 1 && p->build_mode == 4) {
        h.build_kind = BUILD_ROWS;
    } else {
        h.build_kind = BUILD_SLOT;
    }
    const int d = p->dims;
    const int nchunks = (int)((ncell + SC_CHUNK - 1) / SC_CHUNK);
    int rc = SC_OK;
    rc |= A(allocs, &h.pos, n, err);
    rc |= A(allocs, &h.pos_alt, n, err);
    rc |= A(allocs, &h.v, (size_t)n * d, err);
    rc |= A(allocs, &h.v_alt, (size_t)n * d, err);
    rc |= A(allocs, &h.F, (size_t)n * d, err);
    rc |= A(allocs, &h.disp, (size_t)n * d, err);
    rc |= A(allocs, &h.orig, n, err);
    rc |= A(allocs, &h.orig_alt, n, err);
    rc |= A(allocs, &h.cell_of, n, err);
    rc |= A(allocs, &h.rank, n, err);
    rc |= A(allocs, &h.key, n, err);
    rc |= A(allocs, &h.chunk_off, (size_t)nchunks, err);
    rc |= A(allocs, &h.chunk_cnt, (size_t)nchunks * SC_CHUNK_SPLIT, err);
    rc |= A(allocs, &h.cell_cnt, (size_t)ncell, err);
    rc |= A(allocs, &h.cell_start, (size_t)ncell + 1, err);
    rc |= A(allocs, &h.nn, n, err);
    rc |= A(allocs, &h.nbr, (size_t)n * h.cap, err);
    rc |= A(allocs, &h.part, (size_t)h.ntiles * 8, err);
    rc |= A(allocs, &h.stage, (size_t)n * 3, err);
    rc |= A(allocs, &h.dyn, 1, err);
    h.spos = nullptr;
    if (h.prec == SC_PREC_MIXED) rc |= A(allocs, &h.spos, n, err);
    // automatic build mode: the per-slot sweep with the fixed-point prefilter
    // (build mode 2 keeps the FP64-only sweep)
    h.bpos = nullptr;
    if (h.build_kind == BUILD_SLOT && srange >= 1 && p->build_mode == 0) rc |= A(allocs, &h.bpos, n, err);
    if (rc != SC_OK) return SC_ERR_OOM;
    build_prefilter_setup(h);

    std::vector<int> ident(n);
    for (int i = 0; i < n; i++) ident[i] = i;
    std::vector<double4> pos(n);
    for (int i = 0; i < n; i++) pos[i] = make_double4(0.0, 0.0, 0.0, D[i]);
    RepDyn dyn;
    std::memset(&dyn, 0, sizeof(dyn));
    dyn.vscale = 1.0;
    dyn.target = LLONG_MAX;
    bool ok = true;
    ok &= hipMemcpy(h.orig, ident.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(h.pos, pos.data(), sizeof(double4) * n, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemset(h.v, 0, sizeof(double) * n * d) == hipSuccess;
    ok &= hipMemset(h.F, 0, sizeof(double) * n * d) == hipSuccess;
    ok &= hipMemset(h.disp, 0, sizeof(double) * n * d) == hipSuccess;
    ok &= hipMemset(h.cell_cnt, 0, sizeof(int) * ncell) == hipSuccess;
    ok &= hipMemset(h.nn, 0, sizeof(int) * n) == hipSuccess;
    ok &= hipMemset(h.part, 0, sizeof(double) * h.ntiles * 8) == hipSuccess;
    ok &= hipMemset(h.chunk_off, 0, sizeof(int) * nchunks) == hipSuccess;
    ok &= hipMemset(h.chunk_cnt, 0, sizeof(int) * nchunks * SC_CHUNK_SPLIT) == hipSuccess;
    ok &= hipMemcpy(h.dyn, &dyn, sizeof(RepDyn), hipMemcpyHostToDevice) == hipSuccess;
    if (h.spos) {  // the shadow of the initial state (positions 0, diameters D)
        std::vector<uint4> sp(n);
        for (int i = 0; i < n; i++) sp[i] = make_uint4(0u, 0u, 0u, diameter_fixed(D[i], h.dqinv));
        ok &= hipMemcpy(h.spos, sp.data(), sizeof(uint4) * n, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        err = "replica initialisation copy failed";
        return SC_ERR_HIP;
    }
    return SC_OK;
}

void free_all(std::vector<void *> &allocs)
{
    for (void *q : allocs) (void)hipFree(q);
    allocs.clear();
}

}  // namespace sc
