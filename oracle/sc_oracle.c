/*
 * sc_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of SimulationCode.jl's
 * molecular-dynamics hot path (plus, at the end, an all-core OpenMP form of
 * its Newtonian step that serves only as a timing baseline).  It is the checker for the HIP library and the
 * CPU baseline ("port") of the benchmark.  Nothing in the product links it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * No FMA contraction anywhere, so every expression rounds exactly as written.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference checkout).  Julia semantics that matter for rounding:
 *   round(x)      -> rint(x)   (RoundNearest = ties-to-even)
 *   x^2           -> x*x       (literal_pow)
 *   a*b*c         -> (a*b)*c   (n-ary * folds left)
 *   sum(SVector)  -> left fold ((x1+x2)+x3)
 *
 * Parity notes (also in the Python wrapper header):
 *   - CellListMap's internal pair order and d2 are not pinned (the package is
 *     an unversioned dependency absent here); pairs are enumerated here with
 *     calculate_squared_distance and the add_pair! predicates, which is what
 *     every other reference routine uses.
 *   - Rows are filled in ascending (i<j) visiting order; the half-list entry
 *     goes into row min(i,j).  Orientation/order in the reference are
 *     CellListMap internals; compare lists as sets.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_POT_LJ 1
#define OR_POT_BERTHIER 2
#define OR_SYS_NEWTONIAN 0
#define OR_SYS_LANGEVIN 1
#define OR_SYS_BROWNIAN 2

#define OR_PI 3.141592653589793 /* Float64(pi) */

#define OR_OK 0
#define OR_ERR_OVERFLOW -1
#define OR_ERR_ARG -2
#define OR_ERR_NOT_CONVERGED -3

/* Mirrors ParameterStruct + System + InteractionPotential fields the hot path
 * reads (src/DataStructs.jl:10-19, 45-87, 138-153) plus the neighbour-struct
 * scalars (src/DataStructs.jl:243-257, 349-354). */
typedef struct {
    int32_t dims;
    int32_t potential;          /* OR_POT_* */
    int32_t system;             /* OR_SYS_* */
    int32_t rescale_temperature;
    int64_t N;
    int64_t maxneighbours;
    double box_size;
    double force_cutoff2;
    double skin_distance;
    double verlet_sphere_size;  /* skin + force_cutoff*max(D) */
    /* LJ */
    double lj_eps, lj_sigma, lj_Uc;
    /* Berthier */
    double b_c0, b_c2, b_c4, b_zeta;
    /* dynamics */
    double kBT, gamma, dt, m;
} or_params;

/* src/DataStructs.jl:15-18 */
double or_lj_Uc(double eps, double sigma, double rc)
{
    double q = sigma / rc;
    return 4.0 * eps * (pow(q, 12.0) - pow(q, 6.0));
}

/* src/DataStructs.jl:53-58 (n = 12 asserted at :54) */
void or_berthier_constants(double rc, double out[3])
{
    double n = 12.0;
    out[0] = -(1.0 / 8.0) * (8.0 + 6.0 * n + n * n) * pow(rc, -n);
    out[1] = 1.0 / 4.0 * n * (4.0 + n) * pow(rc, -2.0 - n);
    out[2] = -(1.0 / 8.0) * n * (2.0 + n) * pow(rc, -4.0 - n);
}

/* src/AuxiliaryFunctions.jl:311-315 */
double or_squared_distance(const double *ri, const double *rj, int dims, double L)
{
    double s = 0.0;
    for (int c = 0; c < dims; c++) {
        double dr = ri[c] - rj[c];
        dr -= rint(dr / L) * L;
        double sq = dr * dr;
        s = (c == 0) ? sq : s + sq;
    }
    return s;
}

/* find_mean_D -- src/InteractionPotential.jl:1-11 (LJ :5-7, Berthier :9-11) */
static inline double mean_d(const or_params *p, double Di, double Dj)
{
    if (p->potential == OR_POT_LJ)
        return p->lj_sigma * p->lj_sigma;          /* quirk: sigma^2 */
    return (Di + Dj) * 0.5 * (1.0 - p->b_zeta * fabs(Di - Dj));
}

/* potential -- src/InteractionPotential.jl:17-25 (LJ), :87-95 (Berthier) */
static inline double potential(const or_params *p, double r2, double md2)
{
    if (p->potential == OR_POT_LJ) {
        double xi2 = md2 / r2;
        double xi4 = xi2 * xi2;
        double xi6 = xi2 * xi4;
        double xi12 = xi6 * xi6;
        return 4.0 * p->lj_eps * (xi12 - xi6) - p->lj_Uc;
    }
    double invxi2 = r2 / md2;
    double invxi4 = invxi2 * invxi2;
    double xi12 = 1.0 / (invxi4 * invxi4 * invxi4);
    return xi12 + p->b_c4 * invxi4 + p->b_c2 * invxi2 + p->b_c0;
}

/* force -- src/InteractionPotential.jl:39-46 (LJ), :98-107 (Berthier) */
static inline double force_divr(const or_params *p, double r2, double md2)
{
    if (p->potential == OR_POT_LJ) {
        double inv_r2 = 1.0 / r2;
        double xi2 = md2 * inv_r2;
        double xi6 = xi2 * xi2 * xi2;
        double xi12 = xi6 * xi6;
        return p->lj_eps * (48.0 * xi12 - 24.0 * xi6) * inv_r2;
    }
    double inv_md2 = 1.0 / md2;
    double invxi2 = r2 * inv_md2;
    double invxi4 = invxi2 * invxi2;
    double xi14 = 1.0 / (invxi4 * invxi4 * invxi4 * invxi2);
    return -2.0 * inv_md2 * (p->b_c2 + 2.0 * p->b_c4 * invxi2 - 6.0 * xi14);
}

/* ------------------------------------------------------------------ lists */

/* src/NeighborLists.jl:82-104 (add_pair!).  Lists are row-per-particle,
 * 0-based: list[i*maxneigh + k]. */
static int add_pair(const or_params *p, int64_t i, int64_t j, double d2,
                    const double *D, int64_t *nfull, int64_t *full,
                    int64_t *nhalf, int64_t *half)
{
    int64_t M = p->maxneighbours;
    double r_cutoff = sqrt(p->force_cutoff2);          /* NeighborLists.jl:44 */
    if (nfull[i] >= M || nfull[j] >= M)
        return OR_ERR_OVERFLOW;
    full[i * M + nfull[i]++] = j;
    double md = mean_d(p, D[i], D[j]);
    double rh = r_cutoff * md + p->skin_distance;
    if (d2 < rh * rh) {
        if (nhalf[i] >= M)
            return OR_ERR_OVERFLOW;
        half[i * M + nhalf[i]++] = j;
    }
    full[j * M + nfull[j]++] = i;
    return OR_OK;
}

/* src/NeighborLists.jl:24-53 with CellListMap's map_pairwise! restated as an
 * O(N^2) sweep over unordered pairs with d2 <= R_v^2. Zeroes the
 * displacement arrays (NeighborLists.jl:35-42) if they are given. */
int or_build_lists_brute(const or_params *p, const double *r, const double *D,
                         int64_t *nfull, int64_t *full, int64_t *nhalf, int64_t *half,
                         double *disp, double *disp2)
{
    int64_t N = p->N;
    int dims = p->dims;
    double Rv = p->verlet_sphere_size;
    double Rv2 = Rv * Rv;
    memset(nfull, 0, sizeof(int64_t) * N);
    memset(nhalf, 0, sizeof(int64_t) * N);
    if (disp) memset(disp, 0, sizeof(double) * N * dims);
    if (disp2) memset(disp2, 0, sizeof(double) * N);
    for (int64_t i = 0; i < N; i++) {
        for (int64_t j = i + 1; j < N; j++) {
            double d2 = or_squared_distance(r + i * dims, r + j * dims, dims, p->box_size);
            if (d2 <= Rv2) {
                int rc = add_pair(p, i, j, d2, D, nfull, full, nhalf, half);
                if (rc) return rc;
            }
        }
    }
    return OR_OK;
}

/* Same pair set through a serial cell list (cells of side >= R_v, 3^dims
 * stencil); the CPU-baseline path for large N.  Falls back to the O(N^2)
 * sweep when fewer than 3 cells fit per side. */
int or_build_lists_cells(const or_params *p, const double *r, const double *D,
                         int64_t *nfull, int64_t *full, int64_t *nhalf, int64_t *half,
                         double *disp, double *disp2)
{
    int64_t N = p->N;
    int dims = p->dims;
    double L = p->box_size;
    double Rv = p->verlet_sphere_size;
    double Rv2 = Rv * Rv;
    int64_t nc = (int64_t)floor(L / Rv);
    if (nc < 3)
        return or_build_lists_brute(p, r, D, nfull, full, nhalf, half, disp, disp2);
    double cs = L / (double)nc;
    int64_t ncell = nc * nc * (dims == 3 ? nc : 1);
    int64_t *head = (int64_t *)malloc(sizeof(int64_t) * ncell);
    int64_t *next = (int64_t *)malloc(sizeof(int64_t) * N);
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * N);
    if (!head || !next || !cell) { free(head); free(next); free(cell); return OR_ERR_ARG; }
    for (int64_t c = 0; c < ncell; c++) head[c] = -1;
    memset(nfull, 0, sizeof(int64_t) * N);
    memset(nhalf, 0, sizeof(int64_t) * N);
    if (disp) memset(disp, 0, sizeof(double) * N * dims);
    if (disp2) memset(disp2, 0, sizeof(double) * N);
    /* build linked cells; insert in descending i so each chain is ascending */
    for (int64_t i = N - 1; i >= 0; i--) {
        int64_t ci[3] = {0, 0, 0};
        for (int c = 0; c < dims; c++) {
            int64_t k = (int64_t)floor(r[i * dims + c] / cs);
            if (k < 0) k = 0;
            if (k >= nc) k = nc - 1;
            ci[c] = k;
        }
        int64_t id = ci[0] + nc * (ci[1] + nc * ci[2]);
        cell[i] = id;
        next[i] = head[id];
        head[id] = i;
    }
    int rc = OR_OK;
    for (int64_t i = 0; i < N && rc == OR_OK; i++) {
        int64_t id = cell[i];
        int64_t cx = id % nc, cy = (id / nc) % nc, cz = id / (nc * nc);
        int zr = (dims == 3) ? 1 : 0;
        for (int dz = -zr; dz <= zr && rc == OR_OK; dz++)
            for (int dy = -1; dy <= 1 && rc == OR_OK; dy++)
                for (int dx = -1; dx <= 1 && rc == OR_OK; dx++) {
                    int64_t nx = (cx + dx + nc) % nc, ny = (cy + dy + nc) % nc;
                    int64_t nz = (dims == 3) ? (cz + dz + nc) % nc : 0;
                    int64_t nid = nx + nc * (ny + nc * nz);
                    for (int64_t j = head[nid]; j >= 0; j = next[j]) {
                        if (j <= i) continue;
                        double d2 = or_squared_distance(r + i * dims, r + j * dims, dims, L);
                        if (d2 <= Rv2) {
                            rc = add_pair(p, i, j, d2, D, nfull, full, nhalf, half);
                            if (rc) break;
                        }
                    }
                }
    }
    free(head); free(next); free(cell);
    return rc;
}

/* ----------------------------------------------------------------- forces */

/* src/MD.jl:85-137 (Berthier 3D), :140-191 (LJ 3D), :194-238 (Berthier 2D),
 * :291-310 (generic path: 2D LJ).  Half-list gather with Newton-3rd scatter.
 * F must be zeroed by the caller path (or_calculate_new_forces). */
static void forces_pairs(const or_params *p, const double *r, const double *D,
                         const int64_t *nhalf, const int64_t *half, double *F)
{
    int64_t N = p->N, M = p->maxneighbours;
    int dims = p->dims;
    double L = p->box_size, fc2 = p->force_cutoff2;
    for (int64_t i = 0; i < N; i++) {
        int64_t nn = nhalf[i];
        if (nn == 0) continue;                               /* MD.jl:280-283 */
        const double *ri = r + i * dims;
        double Di = D[i];
        double f[3] = {0.0, 0.0, 0.0};
        for (int64_t k = 0; k < nn; k++) {
            int64_t j = half[i * M + k];
            const double *rj = r + j * dims;
            double dr[3] = {0.0, 0.0, 0.0};
            for (int c = 0; c < dims; c++) {
                dr[c] = ri[c] - rj[c];
                dr[c] -= rint(dr[c] / L) * L;
            }
            double rij2 = dr[0] * dr[0] + dr[1] * dr[1];
            if (dims == 3) rij2 = rij2 + dr[2] * dr[2];
            double Fs;
            if (p->potential == OR_POT_BERTHIER) {           /* MD.jl:120-126 */
                double Dj = D[j];
                double md = (Di + Dj) * 0.5 * (1.0 - p->b_zeta * fabs(Di - Dj));
                double md2 = md * md;
                double inv_md2 = 1.0 / md2;
                double invxi2 = rij2 * inv_md2;
                double invxi4 = invxi2 * invxi2;
                double xi14 = 1.0 / (invxi4 * invxi4 * invxi4 * invxi2);
                Fs = (fc2 * md2 < rij2) ? 0.0
                     : -2.0 * inv_md2 * (p->b_c2 + 2.0 * p->b_c4 * invxi2 - 6.0 * xi14);
            } else if (dims == 3) {                          /* MD.jl:175-180 */
                double s2 = p->lj_sigma * p->lj_sigma;
                double xi2 = s2 / rij2;
                double xi6 = xi2 * xi2 * xi2;
                double xi12 = xi6 * xi6;
                Fs = (fc2 < rij2) ? 0.0 : p->lj_eps * (48.0 * xi12 - 24.0 * xi6) / rij2;
            } else {                                         /* MD.jl:299-307 */
                double md = mean_d(p, Di, D[j]);
                double md2 = md * md;
                Fs = (fc2 * md2 > rij2) ? force_divr(p, rij2, md2) : 0.0;
            }
            for (int c = 0; c < dims; c++) {
                f[c] += Fs * dr[c];
                F[j * dims + c] -= Fs * dr[c];
            }
        }
        for (int c = 0; c < dims; c++)
            F[i * dims + c] += f[c];
    }
}

/* src/MD.jl:262-316 calculate_new_forces! followed by add_new_random_force!
 * (src/MD.jl:30-43 Langevin, :67-78 Brownian, :80-82 Newtonian).  `noise`
 * is the N x dims AoS block randn! would have produced (may be NULL for
 * Newtonian). */
void or_calculate_new_forces(const or_params *p, const double *r, const double *D,
                             const int64_t *nhalf, const int64_t *half,
                             const double *v, const double *noise, double *F)
{
    int64_t N = p->N;
    int dims = p->dims;
    memset(F, 0, sizeof(double) * N * dims);
    forces_pairs(p, r, D, nhalf, half, F);
    if (p->system == OR_SYS_NEWTONIAN || noise == NULL)
        return;
    double factor = sqrt(2.0 * p->kBT * p->gamma / p->dt);
    for (int64_t i = 0; i < N; i++)
        for (int c = 0; c < dims; c++) {
            double add;
            if (p->system == OR_SYS_LANGEVIN)
                add = -p->gamma * v[i * dims + c] + factor * noise[i * dims + c];
            else
                add = factor * noise[i * dims + c];
            F[i * dims + c] = F[i * dims + c] + add;
        }
}

/* ---------------------------------------------------------------- energy */

/* src/AuxiliaryFunctions.jl:546-584 (find_energy) with the specialised inner
 * loops :381-413 (Berthier 3D), :416-444 (Berthier 2D), :477-507 (LJ 3D) and
 * the generic loop :568-581 (2D LJ). */
static double find_energy(const or_params *p, int64_t i, const double *r, const double *D,
                          const int64_t *nfull, const int64_t *full)
{
    int64_t nn = nfull[i], M = p->maxneighbours;
    int dims = p->dims;
    double L = p->box_size, fc2 = p->force_cutoff2;
    double e = 0.0;
    if (nn == 0) return 0.0;
    const double *ri = r + i * dims;
    double Di = D[i];
    int generic = (p->potential == OR_POT_LJ && dims == 2);
    for (int64_t k = 0; k < nn; k++) {
        int64_t j = full[i * M + k];
        double rij2 = or_squared_distance(ri, r + j * dims, dims, L);
        double md, md2;
        if (p->potential == OR_POT_LJ && !generic) {
            md = p->lj_sigma;                       /* LJ_inner_loop: mean_d = σ */
            md2 = md * md;
        } else {
            md = mean_d(p, Di, D[j]);
            md2 = md * md;
        }
        if (generic) {
            if (fc2 * md2 > rij2) e += potential(p, rij2, md2);
        } else {
            double v;
            if (p->potential == OR_POT_LJ) {        /* :498-504 */
                double xi2 = md2 / rij2;
                double xi6 = xi2 * xi2 * xi2;
                double xi12 = xi6 * xi6;
                v = 4.0 * p->lj_eps * (xi12 - xi6) - p->lj_Uc;
            } else {
                v = potential(p, rij2, md2);
            }
            e += (fc2 * md2 < rij2) ? 0.0 : v;
        }
    }
    return e;
}

/* src/AuxiliaryFunctions.jl:22-34 */
/* src/AuxiliaryFunctions.jl:186-196 (compute_potential_energy!): E[i] =
 * find_energy(i) / 2. */
void or_potential_energies(const or_params *p, const double *r, const double *D, const int64_t *nfull,
                           const int64_t *full, double *E)
{
    for (int64_t i = 0; i < p->N; i++) E[i] = find_energy(p, i, r, D, nfull, full) / 2.0;
}

double or_full_energy(const or_params *p, const double *r, const double *D,
                      const int64_t *nfull, const int64_t *full)
{
    double e = 0.0;
    for (int64_t i = 0; i < p->N; i++)
        e += find_energy(p, i, r, D, nfull, full);
    return e / 2.0;
}

/* src/AuxiliaryFunctions.jl:86-110 (O(N^2), generic potential()) */
double or_energy_no_neigh(const or_params *p, const double *r, const double *D)
{
    int dims = p->dims;
    double e = 0.0;
    for (int64_t i = 0; i < p->N; i++)
        for (int64_t j = 0; j < p->N; j++) {
            if (i == j) continue;
            double rij2 = or_squared_distance(r + i * dims, r + j * dims, dims, p->box_size);
            double md = mean_d(p, D[i], D[j]);
            double md2 = md * md;
            e += (p->force_cutoff2 * md2 < rij2) ? 0.0 : potential(p, rij2, md2);
        }
    return e / 2.0;
}

/* src/AuxiliaryFunctions.jl:57-65 */
double or_kinetic_energy(const or_params *p, const double *v)
{
    int dims = p->dims;
    double ek = 0.0;
    for (int64_t i = 0; i < p->N; i++) {
        double s = v[i * dims] * v[i * dims];
        for (int c = 1; c < dims; c++) s = s + v[i * dims + c] * v[i * dims + c];
        ek += s;
    }
    return ek * (p->m / 2.0);
}

/* src/MD.jl:391-402 */
void or_rescale_velocities(const or_params *p, double *v)
{
    int dims = p->dims;
    double ek = 0.0;
    for (int64_t i = 0; i < p->N; i++) {
        double s = v[i * dims] * v[i * dims];
        for (int c = 1; c < dims; c++) s = s + v[i * dims + c] * v[i * dims + c];
        ek += s;
    }
    ek *= p->m / 2.0;
    double kT = 2.0 * ek / ((double)dims * (double)p->N);
    double factor = sqrt(p->kBT / kT);
    for (int64_t i = 0; i < p->N * dims; i++) v[i] *= factor;
}

/* ------------------------------------------------------------ time steps */

static int rebuild(const or_params *p, int use_cells, const double *r, const double *D,
                   int64_t *nfull, int64_t *full, int64_t *nhalf, int64_t *half,
                   double *disp, double *disp2, int64_t *rebuilds)
{
    (*rebuilds)++;                                      /* NeighborLists.jl:25 */
    if (use_cells)
        return or_build_lists_cells(p, r, D, nfull, full, nhalf, half, disp, disp2);
    return or_build_lists_brute(p, r, D, nfull, full, nhalf, half, disp, disp2);
}

/* Drift with displacement bookkeeping and floor wrap (src/MD.jl:515-522,
 * and the intended Brownian update :457-466). Returns max(disp2). */
static double drift(const or_params *p, double *r, double *disp, double *disp2,
                    const double *src, double scale)
{
    int dims = p->dims;
    double L = p->box_size;
    double mx = 0.0;
    for (int64_t i = 0; i < p->N; i++) {
        double s = 0.0;
        for (int c = 0; c < dims; c++) {
            double dr = scale * src[i * dims + c];
            disp[i * dims + c] += dr;
            double sq = disp[i * dims + c] * disp[i * dims + c];
            s = (c == 0) ? sq : s + sq;
            double rn = r[i * dims + c] + dr;
            rn -= floor(rn / L) * L;
            r[i * dims + c] = rn;
        }
        disp2[i] = s;
        if (s > mx || i == 0) mx = s;
    }
    return mx;
}

/* Runs `nsteps` calls of do_time_step.
 *   Newtonian/Langevin: src/MD.jl:499-554 (velocity Verlet, COM removal,
 *   optional rescale).  Brownian: the intended semantics of src/MD.jl:429-471
 *   (forces incl. noise, subtract mean force, r += (dt/gamma) F, wrap,
 *   skin trigger), which the reference cannot execute as written.
 * `noise` holds nsteps consecutive N x dims blocks (NULL for Newtonian).
 * `use_cells` selects the cell-list rebuild (same pair set). */
int or_do_time_steps(const or_params *p, int64_t nsteps, int use_cells,
                     double *r, double *v, double *F, const double *D,
                     double *disp, double *disp2,
                     int64_t *nfull, int64_t *full, int64_t *nhalf, int64_t *half,
                     const double *noise, int64_t *rebuilds)
{
    int64_t N = p->N;
    int dims = p->dims;
    int64_t nd = N * dims;
    double dt = p->dt, m = p->m;
    double thr = (p->skin_distance / 2.0) * (p->skin_distance / 2.0);
    int rc;
    for (int64_t s = 0; s < nsteps; s++) {
        const double *xi = noise ? noise + s * nd : NULL;
        if (p->system == OR_SYS_BROWNIAN) {
            or_calculate_new_forces(p, r, D, nhalf, half, v, xi, F);
            double Fm[3] = {0.0, 0.0, 0.0};
            for (int64_t i = 0; i < N; i++)
                for (int c = 0; c < dims; c++) Fm[c] += F[i * dims + c];
            for (int c = 0; c < dims; c++) Fm[c] /= (double)N;
            for (int64_t i = 0; i < N; i++)
                for (int c = 0; c < dims; c++) F[i * dims + c] -= Fm[c];
            double mx = drift(p, r, disp, disp2, F, dt / p->gamma);
            if (mx > thr) {
                rc = rebuild(p, use_cells, r, D, nfull, full, nhalf, half, disp, disp2, rebuilds);
                if (rc) return rc;
            }
            continue;
        }
        for (int64_t i = 0; i < nd; i++)                  /* MD.jl:511-513 */
            v[i] = v[i] + dt * F[i] / (2.0 * m);
        double mx = drift(p, r, disp, disp2, v, dt);      /* MD.jl:515-522 */
        if (mx > thr) {                                   /* MD.jl:523-525 */
            rc = rebuild(p, use_cells, r, D, nfull, full, nhalf, half, disp, disp2, rebuilds);
            if (rc) return rc;
        }
        or_calculate_new_forces(p, r, D, nhalf, half, v, xi, F);   /* MD.jl:526 */
        for (int64_t i = 0; i < nd; i++)                  /* MD.jl:529-531 */
            v[i] = v[i] + dt * F[i] / (2.0 * m);
        double vm[3] = {0.0, 0.0, 0.0};                   /* MD.jl:533-541 */
        for (int64_t i = 0; i < N; i++)
            for (int c = 0; c < dims; c++) vm[c] += v[i * dims + c];
        for (int c = 0; c < dims; c++) vm[c] /= (double)N;
        for (int64_t i = 0; i < N; i++)
            for (int c = 0; c < dims; c++) v[i * dims + c] -= vm[c];
        if (p->rescale_temperature && p->system == OR_SYS_NEWTONIAN)   /* :542-553 */
            or_rescale_velocities(p, v);
    }
    return OR_OK;
}

/* Forced rebuild: update_neighbor_lists! (src/NeighborLists.jl:24-53). */
int or_update_neighbor_lists(const or_params *p, int use_cells, const double *r, const double *D,
                             int64_t *nfull, int64_t *full, int64_t *nhalf, int64_t *half,
                             double *disp, double *disp2, int64_t *rebuilds)
{
    return rebuild(p, use_cells, r, D, nfull, full, nhalf, half, disp, disp2, rebuilds);
}

/* find_inherent_structure -- src/InherentStructure.jl:4-56: capped steepest
 * descent with the pair forces of calculate_new_forces!(arrays, parameters,
 * neighborlist) (src/MD.jl:262-316, no thermostat terms).  Per iteration:
 * v = F dt (:29-31); stepsize = max_i |v_i| (:33-36); v *= max_step/stepsize
 * when stepsize > max_step (:38-40); move with displacement bookkeeping and
 * floor wrap (:41-48); rebuild when max disp2 > (skin/2)^2 (:49-51);
 * error = max(max_i |F_i|, stepsize) (:52-56), the loop ending once
 * error <= tolerance.  Lists are rebuilt and displacements zeroed on entry
 * (:8, :13-14).  Returns OR_OK, or OR_ERR_NOT_CONVERGED once iteration
 * max_iterations+1 would start (:21-23).  r, F, v: AoS N x dims. */
int or_find_inherent_structure(const or_params *p_in, int use_cells, double *r, const double *D, double *F, double *v,
                               double *disp, double *disp2, int64_t *nfull, int64_t *full, int64_t *nhalf,
                               int64_t *half, double tolerance, int64_t max_iterations, double max_step, double dt,
                               int64_t *iterations, int64_t *rebuilds, double *maxforce)
{
    or_params p = *p_in;
    p.system = OR_SYS_NEWTONIAN;          /* pair forces only */
    int64_t N = p.N;
    int dims = p.dims;
    double thr = (p.skin_distance / 2.0) * (p.skin_distance / 2.0);
    int rc = rebuild(&p, use_cells, r, D, nfull, full, nhalf, half, disp, disp2, rebuilds);
    if (rc) return rc;
    for (int64_t i = 0; i < N * dims; i++) disp[i] = 0.0;
    for (int64_t i = 0; i < N; i++) disp2[i] = 0.0;
    double err = tolerance * 2.0;
    int64_t it = 0;
    double fmax_last = 0.0;
    while (err > tolerance) {
        it++;
        if (it > max_iterations) {
            *iterations = it - 1;
            *maxforce = fmax_last;
            return OR_ERR_NOT_CONVERGED;
        }
        or_calculate_new_forces(&p, r, D, nhalf, half, NULL, NULL, F);
        for (int64_t i = 0; i < N * dims; i++) v[i] = F[i] * dt;
        double stepsize = 0.0;
        for (int64_t i = 0; i < N; i++) {
            double s = 0.0;
            for (int c = 0; c < dims; c++) s = (c == 0) ? v[i * dims] * v[i * dims] : s + v[i * dims + c] * v[i * dims + c];
            stepsize = fmax(stepsize, sqrt(s));
        }
        if (stepsize > max_step) {
            double f = max_step / stepsize;
            for (int64_t i = 0; i < N * dims; i++) v[i] *= f;
        }
        double mx = drift(&p, r, disp, disp2, v, 1.0);
        if (mx > thr) {
            rc = rebuild(&p, use_cells, r, D, nfull, full, nhalf, half, disp, disp2, rebuilds);
            if (rc) return rc;
        }
        double fm = 0.0;
        for (int64_t i = 0; i < N; i++) {
            double s = 0.0;
            for (int c = 0; c < dims; c++) s = (c == 0) ? F[i * dims] * F[i * dims] : s + F[i * dims + c] * F[i * dims + c];
            fm = fmax(fm, sqrt(s));
        }
        fmax_last = fm;
        err = fmax(fm, stepsize);
    }
    *iterations = it;
    *maxforce = fmax_last;
    return OR_OK;
}

/* ------------------------------------------------------------------------
 * Log-step observables
 * ------------------------------------------------------------------------ */

/* calculate_F2_self! -- src/AuxiliaryFunctions.jl:968-987: rold and r AoS,
 * x components only, serial sum over particles. */
double or_f2_self(const or_params *p, const double *rold, const double *r, double kx)
{
    int dims = p->dims;
    double F2s = 0.0;
    for (int64_t i = 0; i < p->N; i++) {
        double krold = kx * rold[i * dims];
        double sinkrold = sin(krold), coskrold = cos(krold);
        double kr = kx * r[i * dims];
        double sinkr = sin(kr), coskr = cos(kr);
        F2s += coskrold * coskr + sinkrold * sinkr;
    }
    return F2s / (double)p->N;
}

/* Y_l^m(theta, phi) of SphericalHarmonics.jl's sphericalharmonic(theta, phi;
 * l, m) (an unpinned dependency, Project.toml; its published definition):
 * orthonormal, Condon-Shortley phase, P_l^m(cos theta) built from
 * P_m^m = (-1)^m (2m-1)!! sin^m(theta) with the signed sine, and
 * Y_l^{-m} = (-1)^m conj(Y_l^m). */
void or_ylm(int l, int m, double theta, double phi, double *re, double *im)
{
    int am = m < 0 ? -m : m;
    double x = cos(theta), s = sin(theta);
    double pmm = 1.0;
    for (int k = 1; k <= am; k++) pmm *= -(2.0 * k - 1.0) * s;
    double plm = pmm;
    if (l > am) {
        double p0 = pmm, p1 = x * (2.0 * am + 1.0) * pmm;
        for (int ll = am + 2; ll <= l; ll++) {
            double p2 = ((2.0 * ll - 1.0) * x * p1 - (ll + am - 1.0) * p0) / (double)(ll - am);
            p0 = p1;
            p1 = p2;
        }
        plm = p1;
    }
    double ratio = 1.0;
    for (int q = l - am + 1; q <= l + am; q++) ratio /= (double)q;
    double K = sqrt((2.0 * l + 1.0) / (4.0 * OR_PI) * ratio);
    double yr = K * plm * cos(am * phi), yi = K * plm * sin(am * phi);
    if (m < 0) {
        double sg = (am % 2) ? -1.0 : 1.0;
        yr = sg * yr;
        yi = -sg * yi;
    }
    *re = yr;
    *im = yi;
}

/* calculate_qlm -- src/AuxiliaryFunctions.jl:758-801: mean of Y_l^m over the
 * full-list neighbours of particle i within q_cutoff; returns their number. */
static int64_t qlm_one(const or_params *p, const double *r, const int64_t *nfull, const int64_t *full,
                       int64_t i, int l, int m, double qc2, double *re, double *im)
{
    int64_t M = p->maxneighbours;
    double L = p->box_size;
    double xi = r[i * 3], yi = r[i * 3 + 1], zi = r[i * 3 + 2];
    double qr = 0.0, qi = 0.0;
    int64_t nq = 0;
    for (int64_t k = 0; k < nfull[i]; k++) {
        int64_t j = full[i * M + k];
        double xij = xi - r[j * 3], yij = yi - r[j * 3 + 1], zij = zi - r[j * 3 + 2];
        xij -= rint(xij / L) * L;
        yij -= rint(yij / L) * L;
        zij -= rint(zij / L) * L;
        if (xij * xij + yij * yij + zij * zij > qc2) continue;
        nq++;
        double theta = atan(sqrt(xij * xij + yij * yij) / zij);
        if (xij == 0.0) xij += 1.0000000000000006e-10; /* 10^-10 = inv(10)^10 in Julia */
        double phi = atan(yij / xij);
        double yr, yim;
        or_ylm(l, m, theta, phi, &yr, &yim);
        qr += yr;
        qi += yim;
    }
    if (nq > 0) {
        qr /= (double)nq;
        qi /= (double)nq;
    }
    *re = qr;
    *im = qi;
    return nq;
}

/* calculate_ql -- src/AuxiliaryFunctions.jl:687-711 (3D). */
double or_ql(const or_params *p, const double *r, const int64_t *nfull, const int64_t *full, int l,
             double q_cutoff)
{
    double qc2 = q_cutoff * q_cutoff, avg = 0.0;
    for (int64_t i = 0; i < p->N; i++) {
        double tot = 0.0;
        for (int m = -l; m <= l; m++) {
            double re, im;
            qlm_one(p, r, nfull, full, i, l, m, qc2, &re, &im);
            double a = hypot(re, im);
            tot += a * a;
        }
        avg += sqrt(4.0 * OR_PI / (2.0 * l + 1.0) * tot);
    }
    return avg / (double)p->N;
}

/* calculate_ql_averaged -- src/AuxiliaryFunctions.jl:832-902 (3D). */
double or_ql_averaged(const or_params *p, const double *r, const int64_t *nfull, const int64_t *full, int l,
                      double q_cutoff)
{
    int64_t N = p->N, M = p->maxneighbours, W = 2 * l + 1;
    double L = p->box_size, qc2 = q_cutoff * q_cutoff;
    double *qre = malloc(sizeof(double) * W * N), *qim = malloc(sizeof(double) * W * N);
    int64_t *nqa = malloc(sizeof(int64_t) * W * N);
    if (!qre || !qim || !nqa) {
        free(qre);
        free(qim);
        free(nqa);
        return NAN;
    }
    for (int64_t i = 0; i < N; i++)
        for (int m = -l; m <= l; m++)
            nqa[i * W + m + l] = qlm_one(p, r, nfull, full, i, l, m, qc2, &qre[i * W + m + l], &qim[i * W + m + l]);
    double avg = 0.0;
    for (int64_t i = 0; i < N; i++) {
        double xi = r[i * 3], yi = r[i * 3 + 1], zi = r[i * 3 + 2];
        double tot = 0.0;
        for (int m = -l; m <= l; m++) {
            double sr = qre[i * W + m + l], si = qim[i * W + m + l];
            for (int64_t k = 0; k < nfull[i]; k++) {
                int64_t j = full[i * M + k];
                double xij = xi - r[j * 3], yij = yi - r[j * 3 + 1], zij = zi - r[j * 3 + 2];
                xij -= rint(xij / L) * L;
                yij -= rint(yij / L) * L;
                zij -= rint(zij / L) * L;
                if (xij * xij + yij * yij + zij * zij > qc2) continue;
                sr += qre[j * W + m + l];
                si += qim[j * W + m + l];
            }
            double den = (double)(nqa[i * W + m + l] + 1);
            sr /= den;
            si /= den;
            double a = hypot(sr, si);
            tot += a * a;
        }
        avg += sqrt(tot * 4.0 * OR_PI / (2.0 * l + 1.0));
    }
    free(qre);
    free(qim);
    free(nqa);
    return avg / (double)N;
}

/* ------------------------------------------------------------------------
 * All-core CPU baseline (timing only; not a checker)
 *
 * The Newtonian velocity-Verlet step of or_do_time_steps with its loops over
 * particles split across OpenMP threads, the parallel form SimulationCode.jl
 * would take with threads: a serial linked-cell binning (cells of side >=
 * R_v), a parallel per-particle stencil sweep in which every particle writes
 * only its own row of partners with d2 < (r_c sigma_ij + skin)^2 (both
 * directions, so no thread writes another particle's row), and forces
 * gathered over those rows (each pair evaluated from both sides, no scatter).
 * Sums run in another order than the serial oracle's, so trajectories agree
 * to rounding, not bit for bit.  Full-list rows are counted for the
 * maxneighbours check but not stored.
 * ------------------------------------------------------------------------ */

static int baseline_rebuild(const or_params *p, const double *r, const double *D, int32_t *nshort,
                            int32_t *shortl, double *disp)
{
    int64_t N = p->N, M = p->maxneighbours;
    int dims = p->dims;
    double L = p->box_size, Rv = p->verlet_sphere_size, Rv2 = Rv * Rv;
    double r_cutoff = sqrt(p->force_cutoff2);
    int64_t nc = (int64_t)floor(L / Rv);
    if (nc < 3) nc = 1;
    double cs = L / (double)nc;
    int64_t ncell = nc * nc * (dims == 3 ? nc : 1);
    int64_t *head = (int64_t *)malloc(sizeof(int64_t) * ncell);
    int64_t *next = (int64_t *)malloc(sizeof(int64_t) * N);
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * N);
    if (!head || !next || !cell) { free(head); free(next); free(cell); return OR_ERR_ARG; }
    for (int64_t c = 0; c < ncell; c++) head[c] = -1;
    for (int64_t i = N - 1; i >= 0; i--) {
        int64_t ci[3] = {0, 0, 0};
        for (int c = 0; c < dims; c++) {
            int64_t k = (int64_t)floor(r[i * dims + c] / cs);
            if (k < 0) k = 0;
            if (k >= nc) k = nc - 1;
            ci[c] = k;
        }
        int64_t id = ci[0] + nc * (ci[1] + nc * ci[2]);
        cell[i] = id;
        next[i] = head[id];
        head[id] = i;
    }
    int overflow = 0;
#pragma omp parallel for schedule(static) reduction(| : overflow)
    for (int64_t i = 0; i < N; i++) {
        int64_t id = cell[i];
        int64_t cx = id % nc, cy = (id / nc) % nc, cz = id / (nc * nc);
        int sr = nc == 1 ? 0 : 1;
        int zr = (dims == 3) ? sr : 0;
        int32_t ns = 0, nf = 0;
        for (int dz = -zr; dz <= zr; dz++)
            for (int dy = -sr; dy <= sr; dy++)
                for (int dx = -sr; dx <= sr; dx++) {
                    int64_t nx = (cx + dx + nc) % nc, ny = (cy + dy + nc) % nc;
                    int64_t nz = (dims == 3) ? (cz + dz + nc) % nc : 0;
                    for (int64_t j = head[nx + nc * (ny + nc * nz)]; j >= 0; j = next[j]) {
                        if (j == i) continue;
                        double d2 = or_squared_distance(r + i * dims, r + j * dims, dims, L);
                        if (d2 > Rv2) continue;
                        nf++;
                        double md = mean_d(p, D[i], D[j]);
                        double rh = r_cutoff * md + p->skin_distance;
                        if (d2 < rh * rh) {
                            if (ns < M) shortl[i * M + ns] = (int32_t)j;
                            ns++;
                        }
                    }
                }
        if (nf > M) overflow = 1;
        nshort[i] = ns < M ? ns : (int32_t)M;
        for (int c = 0; c < dims; c++) disp[i * dims + c] = 0.0;
    }
    free(head); free(next); free(cell);
    return overflow ? OR_ERR_OVERFLOW : OR_OK;
}

static void baseline_forces(const or_params *p, const double *r, const double *D, const int32_t *nshort,
                            const int32_t *shortl, double *F)
{
    int64_t N = p->N, M = p->maxneighbours;
    int dims = p->dims;
    double L = p->box_size, fc2 = p->force_cutoff2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
        const double *ri = r + i * dims;
        double Di = D[i];
        double f[3] = {0.0, 0.0, 0.0};
        for (int32_t k = 0; k < nshort[i]; k++) {
            int64_t j = shortl[i * M + k];
            const double *rj = r + j * dims;
            double dr[3] = {0.0, 0.0, 0.0};
            for (int c = 0; c < dims; c++) {
                dr[c] = ri[c] - rj[c];
                dr[c] -= rint(dr[c] / L) * L;
            }
            double rij2 = dr[0] * dr[0] + dr[1] * dr[1];
            if (dims == 3) rij2 = rij2 + dr[2] * dr[2];
            double Fs;
            if (p->potential == OR_POT_BERTHIER) {
                double md = (Di + D[j]) * 0.5 * (1.0 - p->b_zeta * fabs(Di - D[j]));
                double md2 = md * md;
                double inv_md2 = 1.0 / md2;
                double invxi2 = rij2 * inv_md2;
                double invxi4 = invxi2 * invxi2;
                double xi14 = 1.0 / (invxi4 * invxi4 * invxi4 * invxi2);
                Fs = (fc2 * md2 < rij2) ? 0.0 : -2.0 * inv_md2 * (p->b_c2 + 2.0 * p->b_c4 * invxi2 - 6.0 * xi14);
            } else if (dims == 3) {
                double s2 = p->lj_sigma * p->lj_sigma;
                double xi2 = s2 / rij2;
                double xi6 = xi2 * xi2 * xi2;
                double xi12 = xi6 * xi6;
                Fs = (fc2 < rij2) ? 0.0 : p->lj_eps * (48.0 * xi12 - 24.0 * xi6) / rij2;
            } else {
                double md = mean_d(p, Di, D[j]);
                double md2 = md * md;
                Fs = (fc2 * md2 > rij2) ? force_divr(p, rij2, md2) : 0.0;
            }
            for (int c = 0; c < dims; c++) f[c] += Fs * dr[c];
        }
        for (int c = 0; c < dims; c++) F[i * dims + c] = f[c];
    }
}

/* nsteps Newtonian velocity-Verlet steps (src/MD.jl:499-541 operation by
 * operation) on `nthreads` threads.  The lists are rebuilt on entry. */
int or_baseline_steps(const or_params *p, int64_t nsteps, int nthreads, double *r, double *v, double *F,
                      const double *D, int64_t *rebuilds)
{
    if (p->system != OR_SYS_NEWTONIAN || nthreads < 1) return OR_ERR_ARG;
    int64_t N = p->N, M = p->maxneighbours;
    int dims = p->dims;
    double dt = p->dt, m = p->m, L = p->box_size;
    double thr = (p->skin_distance / 2.0) * (p->skin_distance / 2.0);
    int32_t *nshort = (int32_t *)malloc(sizeof(int32_t) * N);
    int32_t *shortl = (int32_t *)malloc(sizeof(int32_t) * N * M);
    double *disp = (double *)malloc(sizeof(double) * N * dims);
    if (!nshort || !shortl || !disp) { free(nshort); free(shortl); free(disp); return OR_ERR_ARG; }
    int prev = omp_get_max_threads();
    omp_set_num_threads(nthreads);
    int rc = baseline_rebuild(p, r, D, nshort, shortl, disp);
    for (int64_t s = 0; s < nsteps && rc == OR_OK; s++) {
        double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
        for (int64_t i = 0; i < N; i++) {
            double s2 = 0.0;
            for (int c = 0; c < dims; c++) {
                int64_t q = i * dims + c;
                v[q] = v[q] + dt * F[q] / (2.0 * m);
                double dr = dt * v[q];
                disp[q] += dr;
                s2 = (c == 0) ? disp[q] * disp[q] : s2 + disp[q] * disp[q];
                double rn = r[q] + dr;
                rn -= floor(rn / L) * L;
                r[q] = rn;
            }
            if (s2 > mx) mx = s2;
        }
        if (mx > thr) {
            (*rebuilds)++;
            rc = baseline_rebuild(p, r, D, nshort, shortl, disp);
            if (rc) break;
        }
        baseline_forces(p, r, D, nshort, shortl, F);
        double vm0 = 0.0, vm1 = 0.0, vm2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : vm0, vm1, vm2)
        for (int64_t i = 0; i < N; i++) {
            for (int c = 0; c < dims; c++) {
                int64_t q = i * dims + c;
                v[q] = v[q] + dt * F[q] / (2.0 * m);
            }
            vm0 += v[i * dims];
            vm1 += v[i * dims + 1];
            if (dims == 3) vm2 += v[i * dims + 2];
        }
        double vm[3] = {vm0 / (double)N, vm1 / (double)N, vm2 / (double)N};
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++)
            for (int c = 0; c < dims; c++) v[i * dims + c] -= vm[c];
    }
    omp_set_num_threads(prev);
    free(nshort); free(shortl); free(disp);
    return rc;
}
