/*
 * sanitize_driver.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Runs every entry point of the oracle (sc_oracle.c, compiled into this
 * translation unit) on small seeded systems so that a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (`make sanitize`) checks its
 * memory accesses and arithmetic: both potentials, 2D and 3D, the three
 * dynamics with injected noise, both list builds, energies, the minimiser,
 * F_s(k), q_l and the all-core baseline step, plus the overflow path.
 * Exits 0 when every call returned what it should; any sanitizer report
 * aborts the run (-fno-sanitize-recover=all).
 */
#include <stdio.h>

#include "sc_oracle.c"

static uint64_t lcg_state = 88172645463325252ull;

static double urand(void) /* uniform [0, 1) from a 64-bit LCG (Knuth's MMIX constants) */
{
    lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg_state >> 11) * 0x1.0p-53;
}

static double nrand(void) /* standard normal, Box-Muller */
{
    double u1 = 1.0 - urand(), u2 = urand();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * OR_PI * u2);
}

static int fails = 0;

static void expect(int ok, const char *what)
{
    if (!ok) {
        fprintf(stderr, "FAILED: %s\n", what);
        fails++;
    }
}

static void run_case(int pot, int dims, int sys)
{
    const int side = dims == 3 ? 5 : 12;
    const int64_t N = dims == 3 ? 125 : 144, M = 200;
    const double rho = pot == OR_POT_LJ ? 1.0 : (dims == 3 ? 1.02 : 1.0);
    const double L = pow((double)N / rho, 1.0 / dims), a = L / side;
    const double rc = pot == OR_POT_LJ ? 2.5 : 1.25;
    or_params p;
    memset(&p, 0, sizeof p);
    p.dims = dims;
    p.potential = pot;
    p.system = sys;
    p.rescale_temperature = sys == OR_SYS_NEWTONIAN;
    p.N = N;
    p.maxneighbours = M;
    p.box_size = L;
    p.force_cutoff2 = rc * rc;
    p.skin_distance = 0.3;
    p.lj_eps = 1.0;
    p.lj_sigma = 1.0;
    p.lj_Uc = or_lj_Uc(1.0, 1.0, rc);
    double cst[3];
    or_berthier_constants(rc, cst);
    p.b_c0 = cst[0];
    p.b_c2 = cst[1];
    p.b_c4 = cst[2];
    p.b_zeta = 0.2;
    p.kBT = 0.2;
    p.gamma = 1.0;
    p.dt = sys == OR_SYS_BROWNIAN ? 0.0002 : 0.002;
    p.m = 1.0;
    const size_t nd = (size_t)N * dims;
    double *r = calloc(nd, sizeof(double)), *v = calloc(nd, sizeof(double)), *F = calloc(nd, sizeof(double));
    double *rold = calloc(nd, sizeof(double)), *D = calloc(N, sizeof(double)), *E = calloc(N, sizeof(double));
    double *disp = calloc(nd, sizeof(double)), *disp2 = calloc(N, sizeof(double));
    double *noise = calloc(nd * 40, sizeof(double));
    int64_t *nfull = calloc(N, sizeof(int64_t)), *full = calloc((size_t)N * M, sizeof(int64_t));
    int64_t *nhalf = calloc(N, sizeof(int64_t)), *half = calloc((size_t)N * M, sizeof(int64_t));
    int64_t *nfull2 = calloc(N, sizeof(int64_t)), *full2 = calloc((size_t)N * M, sizeof(int64_t));
    int64_t *nhalf2 = calloc(N, sizeof(int64_t)), *half2 = calloc((size_t)N * M, sizeof(int64_t));
    double dmax = 0.0;
    for (int64_t i = 0; i < N; i++) {
        int64_t q = i;
        for (int c = 0; c < dims; c++) {
            r[i * dims + c] = ((double)(q % side) + 0.5) * a + 0.1 * (2.0 * urand() - 1.0);
            q /= side;
        }
        D[i] = pot == OR_POT_LJ ? 1.0 : 0.8 + 0.4 * urand();
        dmax = D[i] > dmax ? D[i] : dmax;
        for (int c = 0; c < dims; c++) v[i * dims + c] = 0.3 * nrand();
    }
    p.verlet_sphere_size = p.skin_distance + rc * dmax;
    for (size_t k = 0; k < nd * 40; k++) noise[k] = nrand();
    memcpy(rold, r, nd * sizeof(double));

    expect(or_build_lists_brute(&p, r, D, nfull, full, nhalf, half, disp, disp2) == OR_OK, "brute lists");
    expect(or_build_lists_cells(&p, r, D, nfull2, full2, nhalf2, half2, disp, disp2) == OR_OK, "cell lists");
    int64_t tf = 0, tf2 = 0;
    for (int64_t i = 0; i < N; i++) {
        tf += nfull[i];
        tf2 += nfull2[i];
    }
    expect(tf == tf2 && tf > 0, "brute and cell lists hold as many pairs");
    or_calculate_new_forces(&p, r, D, nhalf, half, v, sys == OR_SYS_NEWTONIAN ? NULL : noise, F);
    const double e_list = or_full_energy(&p, r, D, nfull, full);
    const double e_all = or_energy_no_neigh(&p, r, D);
    expect(fabs(e_list - e_all) <= 1e-10 * (fabs(e_all) + 1.0), "list energy == all-pairs energy");
    or_potential_energies(&p, r, D, nfull, full, E);
    expect(isfinite(or_kinetic_energy(&p, v)), "kinetic energy");
    or_rescale_velocities(&p, v);
    int64_t rebuilds = 0;
    expect(or_update_neighbor_lists(&p, 1, r, D, nfull, full, nhalf, half, disp, disp2, &rebuilds) == OR_OK,
           "forced rebuild");
    expect(or_do_time_steps(&p, 40, 1, r, v, F, D, disp, disp2, nfull, full, nhalf, half,
                            sys == OR_SYS_NEWTONIAN ? NULL : noise, &rebuilds) == OR_OK, "time steps");
    for (size_t k = 0; k < nd; k++) expect(isfinite(r[k]), "finite positions");
    expect(isfinite(or_f2_self(&p, rold, r, 7.2)), "F_s(k)");
    if (dims == 3) {
        for (int l = 1; l <= 8; l++) {
            expect(isfinite(or_ql(&p, r, nfull, full, l, 1.4)), "q_l");
            expect(isfinite(or_ql_averaged(&p, r, nfull, full, l, 1.4)), "averaged q_l");
        }
        double re, im;
        or_ylm(6, -3, -0.4, 2.0, &re, &im);
        expect(isfinite(re) && isfinite(im), "Y_lm");
    }
    int64_t it = 0;
    double fm = 0.0;
    const int mrc = or_find_inherent_structure(&p, 1, r, D, F, v, disp, disp2, nfull, full, nhalf, half, 1e-3,
                                               3000, 0.001, 0.005, &it, &rebuilds, &fm);
    expect((mrc == OR_OK || mrc == OR_ERR_NOT_CONVERGED) && it > 0, "minimiser");
    if (sys == OR_SYS_NEWTONIAN) {
        int64_t nb = 0;
        p.rescale_temperature = 0;
        expect(or_baseline_steps(&p, 5, 2, r, v, F, D, &nb) == OR_OK, "all-core baseline steps");
    }
    /* overflow: rows shorter than the full list */
    p.maxneighbours = 3;
    expect(or_build_lists_cells(&p, r, D, nfull2, full2, nhalf2, half2, disp, disp2) == OR_ERR_OVERFLOW,
           "overflow raised (cells)");
    expect(or_build_lists_brute(&p, r, D, nfull2, full2, nhalf2, half2, disp, disp2) == OR_ERR_OVERFLOW,
           "overflow raised (brute)");
    free(r); free(v); free(F); free(rold); free(D); free(E); free(disp); free(disp2); free(noise);
    free(nfull); free(full); free(nhalf); free(half); free(nfull2); free(full2); free(nhalf2); free(half2);
}

int main(void)
{
    const int pots[2] = {OR_POT_LJ, OR_POT_BERTHIER};
    for (int a = 0; a < 2; a++)
        for (int dims = 2; dims <= 3; dims++)
            for (int sys = 0; sys < 3; sys++) run_case(pots[a], dims, sys);
    double sd = or_squared_distance((const double[3]){0.1, 0.2, 0.3}, (const double[3]){9.9, 0.1, 5.0}, 3, 10.0);
    expect(fabs(sd - (0.2 * 0.2 + 0.1 * 0.1 + 4.7 * 4.7)) < 1e-12, "squared distance (minimum image)");
    if (fails) {
        fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    printf("sanitized oracle run: all entry points ok\n");
    return 0;
}
