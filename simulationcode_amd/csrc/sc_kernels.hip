// sc_kernels.hip -- gfx950 kernels of libsc.
//
// Compiled with -ffp-contract=off: every FP64 expression rounds exactly as
// written, in the operation order of the reference loops it restates, so that
// squared distances (and therefore neighbour-pair sets) are bit-identical to
// the reference's calculate_squared_distance (src/AuxiliaryFunctions.jl:311-315).
//
// Execution model: one thread per particle slot, SC_TB (=256) slots per
// tile, one block per tile (grid = tiles of all replicas in the launch).
// No block ever waits for, or reads the results of, another block of the
// same launch: on gfx950 an agent-scope release fence is an L2 write-back
// (buffer_wbl2), far too costly per block.  Per-replica reductions (COM
// velocity, mean force, max displacement, energies) are written as tile
// partials and finished by k_finish (one block per replica) at the next
// kernel boundary, in a fixed order, so they are deterministic and the
// launch sequence can be replayed without host work.
#include "sc_internal.h"

#include <climits>
#include <math.h>

// Device buffers are reached through the pointers of a Rep read from memory,
// which the compiler cannot place in an address space: it would emit FLAT
// instructions, which also count against the LDS wait counter (an LDS read
// then waits for every outstanding global access).  gp() asserts the global
// address space so every such access is a global_load / global_store.
#define SC_AS1 __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ SC_AS1 T *gp(T *p)
{
    return (SC_AS1 T *)p;
}

#ifndef SC_BUILD_UNROLL
#define SC_BUILD_UNROLL 4  // candidates loaded together per lane in the prefiltered sweep
#endif

// Launch tables (tiles, replicas, rebuild plan) are written only between
// launches, so kernels may read them as constant memory: a read at a
// wavefront-uniform index then becomes a scalar load into SGPRs, also inside
// loops that store to global memory, where the compiler could not otherwise
// rule out that a store changed the table.
#define SC_AS4 __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const SC_AS4 T *cp(const T *p)
{
    return (const SC_AS4 T *)p;
}
// table entry i as a value (word by word: the unused words are never loaded)
template <typename T>
__device__ __forceinline__ T ldc(const T *p, int i)
{
    static_assert(sizeof(T) % 4 == 0, "4-byte words");
    T v;
    const SC_AS4 unsigned *src = (const SC_AS4 unsigned *)(p + i);
    unsigned *dst = reinterpret_cast<unsigned *>(&v);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(T) / 4); w++) dst[w] = src[w];
    return v;
}

// double4 loads / stores through global pointers (the vector type's own
// copy operations take generic references)
typedef double sc_dbl4v __attribute__((ext_vector_type(4)));
typedef unsigned sc_u32x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double4 ld4(const SC_AS1 double4 *p)
{
    const sc_dbl4v v = *(const SC_AS1 sc_dbl4v *)p;
    return make_double4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(SC_AS1 double4 *p, const double4 &x)
{
    *(SC_AS1 sc_dbl4v *)p = sc_dbl4v{x.x, x.y, x.z, x.w};
}

// Relaxed device-scope atomics on global memory (the semantics of HIP's
// atomicAdd / atomicOr).
__device__ __forceinline__ int gatomic_add(SC_AS1 int *p, int v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gatomic_or(SC_AS1 int *p, int v)
{
    (void)__hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

namespace {

constexpr int WAVE = 64;
constexpr int NWAVES = SC_TB / WAVE;

// ------------------------------------------------------------ reductions

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}

__device__ inline double wave_max(double x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, WAVE));
    return x;
}

// Block sum of K values per thread; result valid in all threads.
template <int K>
__device__ inline void block_sum(double (&x)[K])
{
    __shared__ double red[NWAVES][K > 0 ? K : 1];
    int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; k++) x[k] = wave_sum(x[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) red[w][k] = x[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = red[0][k];
#pragma unroll
        for (int q = 1; q < NWAVES; q++) s += red[q][k];
        x[k] = s;
    }
    __syncthreads();
}

__device__ inline double block_max(double x)
{
    __shared__ double red[NWAVES];
    int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    x = wave_max(x);
    __syncthreads();
    if (lane == 0) red[w] = x;
    __syncthreads();
    double m = red[0];
#pragma unroll
    for (int q = 1; q < NWAVES; q++) m = fmax(m, red[q]);
    __syncthreads();
    return m;
}

// Neumaier-compensated pair (sum, compensation) helpers.
__device__ inline void neumaier_add(double &s, double &c, double x)
{
    double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (block b runs on XCD b % 8), each with its own L2.  Tiles are
// cell-sorted, so spatial neighbours are nearby tiles: give every XCD one
// contiguous run of tiles, swept in order, so neighbour gathers hit its L2.
constexpr int NXCD = 8;

__device__ inline int xcd_tile()
{
    const int b = blockIdx.x, nt = gridDim.x;
    if (nt < NXCD) return b;
    const int x = b % NXCD, k = b / NXCD;
    const int base = nt / NXCD, rem = nt % NXCD;
    return x * base + (x < rem ? x : rem) + k;
}

// The rebuild runs when the host requested it or the largest displacement^2
// since the last rebuild exceeds (skin/2)^2 (src/MD.jl:523).  Only that
// decision is needed, so a tile whose own maximum exceeds the threshold sets
// the flag (max(...) > thr  <=>  some tile's max > thr); ordinary steps issue
// no atomics at all.
__device__ inline bool need_rebuild(const Rep &R) { return gp(R.dyn)->rebuild != 0; }

__device__ inline void publish_maxdisp(const Rep &R, double m, int bits = 1)
{
    if (threadIdx.x == 0 && m > R.thr) gatomic_or(&gp(R.dyn)->rebuild, bits);
}

// Per-tile partial slots (8 doubles per tile in Rep::part).
enum { PS_VEC = 0, PS_MAXD2 = 3, PS_AUXMAX = 4, PS_ESUM = 5, PS_ECOMP = 6, PS_SCAL = 7 };

__device__ inline void put_partial(const Rep &R, const Tile &t, int slot, double x)
{
    if (threadIdx.x == 0) gp(R.part)[slot * R.ntiles + t.local] = x;
}

// ------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  Ten rounds, key bumped between rounds.
__host__ __device__ inline void philox4x32_10(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t x0 = in[0], x1 = in[1], x2 = in[2], x3 = in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * x0, p1 = (uint64_t)M1 * x2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
        x0 = y0; x1 = lo1; x2 = y2; x3 = lo0;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

// Two standard normals (Box-Muller) from one Philox block.
__device__ inline void normals2(uint32_t orig, long long step, uint32_t sub, uint64_t seed, double &z0, double &z1)
{
    uint32_t c[4] = {orig, (uint32_t)step, (uint32_t)((unsigned long long)step >> 32), sub};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(c, k, o);
    uint64_t a = ((uint64_t)o[0] << 32) | o[1];
    uint64_t b = ((uint64_t)o[2] << 32) | o[3];
    double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
    double u2 = (double)(b >> 11) * 0x1.0p-53;          // [0, 1)
    double rr = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    z0 = rr * cs;
    z1 = rr * sn;
}

// Four standard normals from one Philox block, Box-Muller in FP32 (the
// mixed form's device noise): 24-bit uniforms u1 in (0, 1], u2 in [0, 1).
// Counter word 3 = 2 keeps the stream apart from the FP64 form's blocks.
__device__ inline void normals4_f32(uint32_t orig, long long step, uint64_t seed, float (&z)[4])
{
    uint32_t c[4] = {orig, (uint32_t)step, (uint32_t)((unsigned long long)step >> 32), 2u};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(c, k, o);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const float u1 = (float)((o[2 * q] >> 8) + 1u) * 0x1.0p-24f;
        const float u2 = (float)(o[2 * q + 1] >> 8) * 0x1.0p-24f;
        const float rr = sqrtf(-2.0f * logf(u1));
        float sn, cs;
        sincospif(2.0f * u2, &sn, &cs);
        z[2 * q] = rr * cs;
        z[2 * q + 1] = rr * sn;
    }
}

// --------------------------------------------------------- pair physics

// Minimum-image component: dx - round(dx/L)*L (src/AuxiliaryFunctions.jl:313).
// rint(dx*invL) equals rint(dx/L) except when dx/L is within an ulp of a
// half-integer, i.e. |dx| ~ L/2, where the wrapped distance is ~L/2 either
// way -- beyond every list radius and cutoff (L > 2 R_v is enforced), so no
// predicate or force can change.
// The shift n*L is exact (n is -1, 0 or 1 for wrapped positions), so the fused
// form rounds once, exactly like dx - t.
__device__ inline double min_image(double dx, double L, double invL)
{
    const double n = rint(dx * invL);
    return fma(-n, L, dx);
}

// FP64 reciprocal: hardware estimate refined by two Newton-Raphson steps
// r <- r + r(1 - x r) (quadratic convergence; result within ~1 ulp of 1/x
// for the normal, finite x met here), in place of a correctly rounded
// division.
__device__ inline double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// FP64 force/r of one pair, the formulas of the reference loops:
// Berthier src/MD.jl:120-126 (3D) / :228-234 (2D); LJ 3D :175-180; generic
// (2D LJ) :299-307 with force() of src/InteractionPotential.jl:39-46.  The
// cutoff tests are evaluated exactly as there; divisions are reciprocals
// (rcp64), which moves the force by a few ulp.
template <int DIMS, int POT>
__device__ inline double pair_fdivr64(const Rep &R, double rij2, double Di, double Dj)
{
    if (POT == 2) {
        double md = (Di + Dj) * 0.5 * (1.0 - R.zeta * fabs(Di - Dj));
        double md2 = md * md;
        double inv_md2 = rcp64(md2);
        double invxi2 = rij2 * inv_md2;
        double invxi4 = invxi2 * invxi2;
        double xi14 = rcp64(invxi4 * invxi4 * invxi4 * invxi2);
        return (R.fc2 * md2 < rij2) ? 0.0 : -2.0 * inv_md2 * (R.c2 + 2.0 * R.c4 * invxi2 - 6.0 * xi14);
    } else if (DIMS == 3) {
        double inv_r2 = rcp64(rij2);
        double xi2 = R.lj_s2 * inv_r2;
        double xi6 = xi2 * xi2 * xi2;
        double xi12 = xi6 * xi6;
        return (R.fc2 < rij2) ? 0.0 : R.eps * (48.0 * xi12 - 24.0 * xi6) * inv_r2;
    } else {
        double md2 = R.lj_md * R.lj_md;
        if (R.fc2 * md2 > rij2) {
            double inv_r2 = rcp64(rij2);
            double xi2 = md2 * inv_r2;
            double xi6 = xi2 * xi2 * xi2;
            double xi12 = xi6 * xi6;
            return R.eps * (48.0 * xi12 - 24.0 * xi6) * inv_r2;
        }
        return 0.0;
    }
}

// Reciprocal for the mixed form: the FP32 hardware estimate (relative error
// ~2^-23) refined by one FP64 Newton step r <- r + r (1 - x r), which squares
// the error (~1e-14), for the normal, finite x met here.
__device__ inline double rcp_mixed(double x)
{
    const double r = (double)__builtin_amdgcn_rcpf((float)x);
    return fma(r, fma(-x, r, 1.0), r);
}

// Mixed form, one pair from the fixed-point records qi, qj: separation from
// the wrapping differences of the 32-bit coordinates (the minimum image,
// exact), products and sums in FP64 (no slower than FP32 on this hardware
// without packed FP32), the reciprocals from FP32 estimates.  The force law
// and cutoff tests are those of pair_fdivr64 (src/MD.jl:120-126, 175-180,
// 228-234, 299-307).  Accumulates force/r times the separation into f.
template <int DIMS, int POT>
__device__ inline void pair_mixed(const Rep &R, const uint4 qi, const double Di, const uint4 qj, double (&f)[3])
{
    const double dx = (double)(int)(qi.x - qj.x) * R.lq;
    const double dy = (double)(int)(qi.y - qj.y) * R.lq;
    const double dz = DIMS == 3 ? (double)(int)(qi.z - qj.z) * R.lq : 0.0;
    double r2 = fma(dy, dy, dx * dx);
    if (DIMS == 3) r2 = fma(dz, dz, r2);
    double fs;
    if (POT == 2) {
        const double Dj = (double)qj.w * R.dq;
        const double md = (Di + Dj) * fma(-0.5 * R.zeta, fabs(Di - Dj), 0.5);
        const double md2 = md * md;
        const double inv_md2 = rcp_mixed(md2);
        const double x = r2 * inv_md2;  // (r / sigma_ij)^2
        const double x2 = x * x, x4 = x2 * x2;
        const double xi14 = rcp_mixed(x4 * x2 * x);
        fs = (R.fc2 * md2 < r2) ? 0.0 : -2.0 * inv_md2 * fma(-6.0, xi14, fma(2.0 * R.c4, x, R.c2));
    } else if (DIMS == 3) {
        const double inv_r2 = rcp_mixed(r2);
        const double xi2 = R.lj_s2 * inv_r2;
        const double xi6 = xi2 * xi2 * xi2;
        fs = (R.fc2 < r2) ? 0.0 : R.eps * fma(48.0, xi6 * xi6, -24.0 * xi6) * inv_r2;
    } else {
        const double md2 = R.lj_md * R.lj_md;
        fs = 0.0;
        if (R.fc2 * md2 > r2) {
            const double inv_r2 = rcp_mixed(r2);
            const double xi2 = md2 * inv_r2;
            const double xi6 = xi2 * xi2 * xi2;
            fs = R.eps * fma(48.0, xi6 * xi6, -24.0 * xi6) * inv_r2;
        }
    }
    f[0] = fma(fs, dx, f[0]);
    f[1] = fma(fs, dy, f[1]);
    if (DIMS == 3) f[2] = fma(fs, dz, f[2]);
}

__device__ __forceinline__ uint4 ldq(const uint4 *p)
{
    const sc_u32x4v v = *(const SC_AS1 sc_u32x4v *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Pair energy with the cutoff test of find_energy's inner loops
// (src/AuxiliaryFunctions.jl:381-413, :416-444, :477-507; generic :568-581).
template <int DIMS, int POT>
__device__ inline double pair_energy(const Rep &R, double rij2, double Di, double Dj)
{
    if (POT == 2) {
        double md = (Di + Dj) * 0.5 * (1.0 - R.zeta * fabs(Di - Dj));
        double md2 = md * md;
        double invxi2 = rij2 / md2;
        double invxi4 = invxi2 * invxi2;
        double xi12 = 1.0 / (invxi4 * invxi4 * invxi4);
        double e = xi12 + R.c4 * invxi4 + R.c2 * invxi2 + R.c0;
        return (R.fc2 * md2 < rij2) ? 0.0 : e;
    } else if (DIMS == 3) {
        double md2 = R.lj_s2;  // sigma * sigma
        double xi2 = md2 / rij2;
        double xi6 = xi2 * xi2 * xi2;
        double xi12 = xi6 * xi6;
        double e = 4.0 * R.eps * (xi12 - xi6) - R.Uc;
        return (R.fc2 * md2 < rij2) ? 0.0 : e;
    } else {
        double md2 = R.lj_md * R.lj_md;
        if (R.fc2 * md2 > rij2) {
            double xi2 = md2 / rij2;
            double xi4 = xi2 * xi2;
            double xi6 = xi2 * xi4;
            double xi12 = xi6 * xi6;
            return 4.0 * R.eps * (xi12 - xi6) - R.Uc;
        }
        return 0.0;
    }
}

// Potential of calculate_full_energy_no_neigh (generic potential(),
// src/AuxiliaryFunctions.jl:94-108 with potential() of
// src/InteractionPotential.jl:17-25 (LJ) and :87-95 (Berthier)): mean_d from
// find_mean_D (:1-11) for every potential.
template <int POT>
__device__ inline double pair_energy_generic(const Rep &R, double rij2, double Di, double Dj)
{
    double md, md2;
    if (POT == 2) {
        md = (Di + Dj) * 0.5 * (1.0 - R.zeta * fabs(Di - Dj));
        md2 = md * md;
        double invxi2 = rij2 / md2;
        double invxi4 = invxi2 * invxi2;
        double xi12 = 1.0 / (invxi4 * invxi4 * invxi4);
        double e = xi12 + R.c4 * invxi4 + R.c2 * invxi2 + R.c0;
        return (R.fc2 * md2 < rij2) ? 0.0 : e;
    }
    md = R.lj_md;
    md2 = md * md;
    double xi2 = md2 / rij2;
    double xi4 = xi2 * xi2;
    double xi6 = xi2 * xi4;
    double xi12 = xi6 * xi6;
    double e = 4.0 * R.eps * (xi12 - xi6) - R.Uc;
    return (R.fc2 * md2 < rij2) ? 0.0 : e;
}

// Squared minimum-image distance, left fold (src/AuxiliaryFunctions.jl:311-315).
template <int DIMS>
__device__ inline double sqdist(const Rep &R, double4 a, double4 b, double &dx, double &dy, double &dz)
{
    dx = min_image(a.x - b.x, R.L, R.invL);
    dy = min_image(a.y - b.y, R.L, R.invL);
    double d2 = dx * dx + dy * dy;
    if (DIMS == 3) {
        dz = min_image(a.z - b.z, R.L, R.invL);
        d2 = d2 + dz * dz;
    } else {
        dz = 0.0;
    }
    return d2;
}

// Cell coordinate of a wrapped position component.
__device__ inline int cell_coord(double x, const Rep &R)
{
    int c = (int)floor(x * R.cinv);
    c = c < 0 ? 0 : c;
    return c >= R.nc ? R.nc - 1 : c;
}

template <int DIMS>
__device__ inline void cell_of_pos(const Rep &R, double4 p, int (&cc)[3])
{
    cc[0] = cell_coord(p.x, R);
    cc[1] = cell_coord(p.y, R);
    cc[2] = DIMS == 3 ? cell_coord(p.z, R) : 0;
}

__device__ inline int wrapc(int c, int nc) { return c < 0 ? c + nc : (c >= nc ? c - nc : c); }

// Mixed precision: the 32-bit fixed-point record of slot k, written wherever
// pos is.  A coordinate maps to u = rint(x * 2^32 / L) mod 2^32 (modular, so a
// position a rounding outside [0, L) still lands on its periodic image); the
// diameter to w = rint(D * dqinv), exact to the scale's resolution.
__device__ inline unsigned fixed_coord(double x, double uq) { return (unsigned)(long long)rint(x * uq); }

__device__ inline void put_shadow(const Rep &R, int k, const double4 p)
{
    if (!R.spos) return;
    *(SC_AS1 sc_u32x4v *)(R.spos + k) =
        sc_u32x4v{fixed_coord(p.x, R.uq), fixed_coord(p.y, R.uq), R.dims == 3 ? fixed_coord(p.z, R.uq) : 0u,
                  (unsigned)(long long)rint(p.w * R.dqinv)};
}

// ====================================================================
// Integrator kernels
// ====================================================================

// First half-kick + drift + displacement bookkeeping + floor wrap + skin
// trigger: src/MD.jl:511-525.  Applies the pending COM shift / rescale of
// the previous step first (src/MD.jl:533-553), in the reference's order.
template <int DIMS>
__device__ void kick_drift_tile(const Rep &R, const Tile &t)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int n = R.n;
    const int i = t.first + threadIdx.x;
    const int pend = dyn->pending;
    double s2 = 0.0;
    if (i < n) {
        double4 p = ld4(gp(R.pos) + (i));
        double pc[3] = {p.x, p.y, p.z};
        double s = 0.0;
        bool fin = true;
#pragma unroll
        for (int c = 0; c < DIMS; c++) {
            double vc = gp(R.v)[c * n + i];
            if (pend & 1) vc = vc - dyn->vmean[c];
            if (pend & 2) vc = vc * dyn->vscale;
            vc = vc + R.dt * gp(R.F)[c * n + i] / (2.0 * R.m);
            double dr = R.dt * vc;
            double dc = gp(R.disp)[c * n + i] + dr;
            s = (c == 0) ? dc * dc : s + dc * dc;
            double rn = pc[c] + dr;
            rn -= floor(rn / R.L) * R.L;
            pc[c] = rn;
            fin = fin && isfinite(rn);
            gp(R.v)[c * n + i] = vc;
            gp(R.disp)[c * n + i] = dc;
        }
        const double4 pn = make_double4(pc[0], pc[1], pc[2], p.w);
        st4(gp(R.pos) + (i), pn);
        put_shadow(R, i, pn);
        s2 = s;
        if (!fin) gatomic_or(&dyn->err, SC_EBIT_NONFINITE);
    }
    publish_maxdisp(R, block_max(s2));
}

__global__ __launch_bounds__(SC_TB) void k_kick_drift(const Tile *tiles, const Rep *reps, int ahead)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (R.sys == 2) return;
    // halt (and step) are written only by the force finish, never during this launch
    if (ahead && (gp(R.dyn)->halt || gp(R.dyn)->step >= gp(R.dyn)->target)) return;
    if (R.dims == 3)
        kick_drift_tile<3>(R, t);
    else
        kick_drift_tile<2>(R, t);
}

// Brownian drift: F -= mean(F); r += (dt/gamma) F; bookkeeping; wrap;
// trigger.  The intended semantics of src/MD.jl:445-470.
template <int DIMS>
__device__ void drift_brownian_tile(const Rep &R, const Tile &t)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int n = R.n;
    const int i = t.first + threadIdx.x;
    double s2 = 0.0;
    if (i < n) {
        double4 p = ld4(gp(R.pos) + (i));
        double pc[3] = {p.x, p.y, p.z};
        double factor = R.dt / R.gamma;
        double s = 0.0;
        bool fin = true;
#pragma unroll
        for (int c = 0; c < DIMS; c++) {
            double fc = gp(R.F)[c * n + i] - dyn->Fmean[c];
            double dr = factor * fc;
            double dc = gp(R.disp)[c * n + i] + dr;
            s = (c == 0) ? dc * dc : s + dc * dc;
            double rn = pc[c] + dr;
            rn -= floor(rn / R.L) * R.L;
            pc[c] = rn;
            fin = fin && isfinite(rn);
            gp(R.F)[c * n + i] = fc;
            gp(R.disp)[c * n + i] = dc;
        }
        const double4 pn = make_double4(pc[0], pc[1], pc[2], p.w);
        st4(gp(R.pos) + (i), pn);
        put_shadow(R, i, pn);
        s2 = s;
        if (!fin) gatomic_or(&dyn->err, SC_EBIT_NONFINITE);
    }
    publish_maxdisp(R, block_max(s2));
}

__global__ __launch_bounds__(SC_TB) void k_drift_brownian(const Tile *tiles, const Rep *reps, int ahead)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (R.sys != 2) return;
    // halt is written only by the force finish: bit 0 the chunk halted, bit 1
    // the replica has done its steps (this step's force finish saw it)
    if (ahead && gp(R.dyn)->halt) return;
    if (R.dims == 3)
        drift_brownian_tile<3>(R, t);
    else
        drift_brownian_tile<2>(R, t);
}

// Pair forces over the symmetric short list (gather only, no atomics, so the
// result is bitwise repeatable), then add_new_random_force! and, in MODE 1,
// the second half-kick with COM partial sums.
//   MODE 0: calculate_new_forces! only (src/MD.jl:262-316)
//   MODE 1: + v += dt F/(2m) and COM velocity (src/MD.jl:526-541)
//   MODE 2: Brownian: + mean-force partials (src/MD.jl:442-454)
template <int DIMS, int POT, int PREC, int MODE>
__device__ void force_tile(const Rep &R, const Tile &t)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int n = R.n;
    const int i = t.first + threadIdx.x;
    const long long step = dyn->step;
    double f[3] = {0.0, 0.0, 0.0};
    double red[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double4 pi = ld4(gp(R.pos) + (i));
        const int cnt = gp(R.nn)[i];
        const SC_AS1 int *nb = gp(R.nbr) + i;
        if (PREC == 0) {
            // gathers issued four at a time so several are in flight per
            // lane; accumulation order stays k = 0, 1, 2, ...
            int k = 0;
            for (; k + 4 <= cnt; k += 4) {
                int jj[4];
                double4 pp[4];
#pragma unroll
                for (int u = 0; u < 4; u++) jj[u] = nb[(size_t)(k + u) * n];
#pragma unroll
                for (int u = 0; u < 4; u++) pp[u] = ld4(gp(R.pos) + (jj[u]));
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    double dx, dy, dz;
                    double rij2 = sqdist<DIMS>(R, pi, pp[u], dx, dy, dz);
                    double fs = pair_fdivr64<DIMS, POT>(R, rij2, pi.w, pp[u].w);
                    f[0] += fs * dx;
                    f[1] += fs * dy;
                    if (DIMS == 3) f[2] += fs * dz;
                }
            }
            if (k < cnt) {
                // the last cnt - k < 4 entries, their loads issued together
                // (index i stands in for the absent ones and is not used)
                int jj[3];
                double4 pp[3];
#pragma unroll
                for (int u = 0; u < 3; u++) jj[u] = k + u < cnt ? nb[(size_t)(k + u) * n] : i;
#pragma unroll
                for (int u = 0; u < 3; u++) pp[u] = ld4(gp(R.pos) + (jj[u]));
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    if (k + u >= cnt) break;
                    double dx, dy, dz;
                    double rij2 = sqdist<DIMS>(R, pi, pp[u], dx, dy, dz);
                    double fs = pair_fdivr64<DIMS, POT>(R, rij2, pi.w, pp[u].w);
                    f[0] += fs * dx;
                    f[1] += fs * dy;
                    if (DIMS == 3) f[2] += fs * dz;
                }
            }
        } else {
            // mixed: 16-byte fixed-point records (one load per pair)
            const uint4 qi = ldq(R.spos + i);
            const double Di = (double)qi.w * R.dq;
            int k = 0;
            for (; k + 4 <= cnt; k += 4) {
                int jj[4];
                uint4 qq[4];
#pragma unroll
                for (int u = 0; u < 4; u++) jj[u] = nb[(size_t)(k + u) * n];
#pragma unroll
                for (int u = 0; u < 4; u++) qq[u] = ldq(R.spos + jj[u]);
#pragma unroll
                for (int u = 0; u < 4; u++) pair_mixed<DIMS, POT>(R, qi, Di, qq[u], f);
            }
            if (k < cnt) {
                int jj[3];
                uint4 qq[3];
#pragma unroll
                for (int u = 0; u < 3; u++) jj[u] = k + u < cnt ? nb[(size_t)(k + u) * n] : i;
#pragma unroll
                for (int u = 0; u < 3; u++) qq[u] = ldq(R.spos + jj[u]);
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    if (k + u >= cnt) break;
                    pair_mixed<DIMS, POT>(R, qi, Di, qq[u], f);
                }
            }
        }
        // add_new_random_force! (src/MD.jl:30-43 Langevin, :67-78 Brownian)
        if (R.sys != 0) {
            double xi[3];
            if (R.noise) {
                // injected blocks cover noise_step0 .. noise_step0 + noise_blocks - 1;
                // outside them (a call sequence the host did not plan) no block is read
                const long long blk = step - R.noise_step0;
                if (blk >= 0 && blk < R.noise_blocks) {
                    const SC_AS1 double *nz = gp(R.noise) + (size_t)blk * (size_t)n * DIMS +
                                              (size_t)gp(R.orig)[i] * DIMS;
#pragma unroll
                    for (int c = 0; c < DIMS; c++) xi[c] = nz[c];
                } else {
#pragma unroll
                    for (int c = 0; c < DIMS; c++) xi[c] = 0.0;
                    gatomic_or(&dyn->err, SC_EBIT_NOISE);
                }
            } else if (PREC == 1) {
                float z[4];
                normals4_f32((uint32_t)gp(R.orig)[i], step, R.seed, z);
#pragma unroll
                for (int c = 0; c < DIMS; c++) xi[c] = (double)z[c];
            } else {
                double z0, z1, z2, z3;
                normals2((uint32_t)gp(R.orig)[i], step, 0u, R.seed, z0, z1);
                xi[0] = z0;
                xi[1] = z1;
                if (DIMS == 3) {
                    normals2((uint32_t)gp(R.orig)[i], step, 1u, R.seed, z2, z3);
                    xi[2] = z2;
                }
            }
            double factor = sqrt(2.0 * R.kBT * R.gamma / R.dt);
#pragma unroll
            for (int c = 0; c < DIMS; c++) {
                double add = (R.sys == 1) ? -R.gamma * gp(R.v)[c * n + i] + factor * xi[c] : factor * xi[c];
                f[c] = f[c] + add;
            }
        }
#pragma unroll
        for (int c = 0; c < DIMS; c++) gp(R.F)[c * n + i] = f[c];
        if (MODE == 1) {
#pragma unroll
            for (int c = 0; c < DIMS; c++) {
                double vc = gp(R.v)[c * n + i] + R.dt * f[c] / (2.0 * R.m);
                gp(R.v)[c * n + i] = vc;
                red[c] = vc;
            }
        } else if (MODE == 2) {
#pragma unroll
            for (int c = 0; c < DIMS; c++) red[c] = f[c];
        }
    }
    if (MODE != 0) {
        block_sum<3>(red);
        if (threadIdx.x == 0)
#pragma unroll
            for (int c = 0; c < 3; c++) gp(R.part)[(PS_VEC + c) * R.ntiles + t.local] = red[c];
    }
}

template <int MODE>
__global__ __launch_bounds__(SC_TB) void k_force(const Tile *tiles, const Rep *reps, int sysmask, int ahead)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (!((sysmask >> R.sys) & 1)) return;
    // launched ahead: this step's drift fired the trigger (the rebuild must
    // come first), an earlier one did (the chunk has halted) or the replica
    // has done its steps; ahead == 2: the resumed step of a halted replica only
    if (ahead == 1 && (gp(R.dyn)->halt || gp(R.dyn)->rebuild || gp(R.dyn)->step >= gp(R.dyn)->target)) return;
    if (ahead == 2 && !(gp(R.dyn)->halt & 1)) return;
    if (R.min_tol >= 0.0 && gp(R.dyn)->min_state != 0) return;  // minimiser converged
    const int key = (R.dims == 3 ? 4 : 0) + (R.pot == 2 ? 2 : 0) + (R.prec ? 1 : 0);
    switch (key) {
    case 0: force_tile<2, 1, 0, MODE>(R, t); break;
    case 1: force_tile<2, 1, 1, MODE>(R, t); break;
    case 2: force_tile<2, 2, 0, MODE>(R, t); break;
    case 3: force_tile<2, 2, 1, MODE>(R, t); break;
    case 4: force_tile<3, 1, 0, MODE>(R, t); break;
    case 5: force_tile<3, 1, 1, MODE>(R, t); break;
    case 6: force_tile<3, 2, 0, MODE>(R, t); break;
    default: force_tile<3, 2, 1, MODE>(R, t); break;
    }
}

// Kinetic energy of the COM-shifted velocities -> rescale factor
// (src/MD.jl:542-553).  Runs only for replicas with rescale_temperature.
template <int DIMS>
__device__ void ekin_com_tile(const Rep &R, const Tile &t)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int n = R.n, i = t.first + threadIdx.x;
    double e[1] = {0.0};
    if (i < n) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < DIMS; c++) {
            double vc = gp(R.v)[c * n + i] - dyn->vmean[c];
            s = (c == 0) ? vc * vc : s + vc * vc;
        }
        e[0] = s;
    }
    block_sum<1>(e);
    put_partial(R, t, PS_SCAL, e[0]);
}

__global__ __launch_bounds__(SC_TB) void k_ekin_com(const Tile *tiles, const Rep *reps)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (!R.rescale || R.sys != 0) return;
    if (R.dims == 3)
        ekin_com_tile<3>(R, t);
    else
        ekin_com_tile<2>(R, t);
}

// Apply the pending COM shift / rescale so the state is reference-consistent
// between API calls.
__global__ __launch_bounds__(SC_TB) void k_finalize(const Tile *tiles, const Rep *reps)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int pend = dyn->pending;
    const int n = R.n, i = t.first + threadIdx.x;
    if (pend && i < n) {
        for (int c = 0; c < R.dims; c++) {
            double vc = gp(R.v)[c * n + i];
            if (pend & 1) vc = vc - dyn->vmean[c];
            if (pend & 2) vc = vc * dyn->vscale;
            gp(R.v)[c * n + i] = vc;
        }
    }
}

// ====================================================================
// Neighbour-list rebuild (src/NeighborLists.jl:24-104): counting sort of the
// slots by cell, then a stencil sweep writing the symmetric short list.
// Every kernel returns at once for replicas whose rebuild flag is clear.
// ====================================================================

// A rebuild requested between steps (F stays live).  It also clears the
// replica's error bits: the rebuild validates the lists again (overflow) and
// the next drift the positions (non-finite).
__global__ void k_request_rebuild(Rep *reps, int nreps)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nreps) {
        gp(reps[r].dyn)->rebuild = 1 | REBUILD_KEEP_F;
        gp(reps[r].dyn)->err = 0;
    }
}

// Error bits of every replica into one array (one host copy per check).
// what = 0: error bits; 1: minimiser stop state
__global__ void k_gather_err(const Rep *reps, int nreps, int *out, int what)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nreps) {
        const SC_AS1 RepDyn *d = gp(reps[r].dyn);
        if (what == 2) {
            const long long left = d->target - d->step;
            const int l = left <= 0 ? 0 : (left > (1 << 28) ? (1 << 28) : (int)left);
            out[r] = (d->halt & 1) | (d->rebuild ? 2 : 0) | (l << 2);
        } else {
            out[r] = what == 0 ? d->err : d->min_state;
        }
    }
}

// Rebuild work list.  plan[0] = number of tiles to rebuild, plan[1] = number
// of replicas to rebuild, plan[2] = number of windowed-build runs, then from
// plan[4]: those replicas (ascending), their tiles (ascending) and the
// (replica, run) pairs of their windowed builds.  The rebuild kernels sweep
// these lists only, so a step in which a few replicas of a large batch
// rebuild spreads just their work over the whole grid.
__global__ __launch_bounds__(SC_TB) void k_rebuild_plan(const Rep *reps, int nreps, int ntiles, int *plan)
{
    __shared__ int st[SC_TB], sr[SC_TB], s0[SC_TB], sw[SC_TB];
    __shared__ int base_t, base_r, base_w;
    const int tid = threadIdx.x;
    int *rlist = plan + 4, *tlist = plan + 4 + nreps, *wlist = plan + 4 + nreps + ntiles;
    if (tid == 0) base_t = base_r = base_w = 0;
    for (int b = 0; b < nreps; b += SC_TB) {
        const int r = b + tid;
        int flag = 0, nt = 0, t0 = 0, nw = 0;
        if (r < nreps) {
            const Rep &R = reps[r];
            t0 = R.tile0;
            nt = R.ntiles;
            flag = gp(R.dyn)->rebuild != 0 ? 1 : 0;
            nt = flag ? nt : 0;
            nw = (flag && R.build_kind == BUILD_WIN) ? R.nruns : 0;
        }
        st[tid] = nt;
        sr[tid] = flag;
        s0[tid] = t0;
        sw[tid] = nw;
        __syncthreads();
        for (int o = 1; o < SC_TB; o <<= 1) {  // inclusive scans of tile, replica and run counts
            const int xt = tid >= o ? st[tid - o] : 0, xr = tid >= o ? sr[tid - o] : 0;
            const int xw = tid >= o ? sw[tid - o] : 0;
            __syncthreads();
            st[tid] += xt;
            sr[tid] += xr;
            sw[tid] += xw;
            __syncthreads();
        }
        if (flag) rlist[base_r + sr[tid] - 1] = r;
        const int total = st[SC_TB - 1];
        for (int x = tid; x < total; x += SC_TB) {  // x-th planned tile of this chunk
            int lo = 0, hi = SC_TB - 1;               // first q with st[q] > x
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (st[mid] > x)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            const int before = lo ? st[lo - 1] : 0;
            tlist[base_t + x] = s0[lo] + (x - before);
        }
        const int totw = sw[SC_TB - 1];
        for (int x = tid; x < totw; x += SC_TB) {  // x-th planned run of this chunk
            int lo = 0, hi = SC_TB - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (sw[mid] > x)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            const int before = lo ? sw[lo - 1] : 0;
            wlist[2 * (base_w + x)] = b + lo;
            wlist[2 * (base_w + x) + 1] = x - before;
        }
        __syncthreads();
        if (tid == 0) {
            base_t += total;
            base_r += sr[SC_TB - 1];
            base_w += totw;
        }
        __syncthreads();
    }
    if (tid == 0) {
        plan[0] = base_t;
        plan[1] = base_r;
        plan[2] = base_w;
        plan[3] = 0;
    }
}

// The tiles a rebuild kernel sweeps: the planned list (batches), or, with
// no plan (a single replica), all of the replica's tiles when its flag is set.
struct RebuildWork {
    const int *tl;
    int count;
    __device__ int tile(int w) const { return tl ? cp(tl)[w] : w; }
};

// flag: the single replica's rebuild flag, passed directly so that a step
// without a rebuild costs each kernel one load.
__device__ inline RebuildWork rebuild_work(const int *plan, int nreps, const int *flag, int ntiles)
{
    if (plan) return RebuildWork{plan + 4 + nreps, plan[0]};
    return RebuildWork{nullptr, *gp(flag) != 0 ? ntiles : 0};
}

// Replica of the idx-th rebuilding replica, or -1.
__device__ inline int rebuild_rep(const int *plan, const int *flag, int idx)
{
    if (plan) return idx < plan[1] ? plan[4 + idx] : -1;
    return (idx == 0 && *gp(flag) != 0) ? 0 : -1;
}

// Slot range start of cell c, 0 <= c <= ncell (the scan leaves absolute
// starts and the sentinel cell_start[ncell] = n).
__device__ inline int cstart(const Rep &R, int c)
{
    return gp(R.cell_start)[min(c, R.ncell)];
}

// XCD-aware grid-stride iteration for capped grids: the blocks resident on
// one XCD sweep one contiguous range of work items, in order.
struct XcdRange {
    int cur, end, step;
};

__device__ inline XcdRange xcd_range(int total)
{
    const int b = blockIdx.x, g = gridDim.x;
    if (g < NXCD) return XcdRange{b, total, g};
    const int x = b % NXCD, kb = b / NXCD;
    const int nbx = (g - x + NXCD - 1) / NXCD;  // blocks on this XCD
    const int base = total / NXCD, rem = total % NXCD;
    const int c0 = x * base + (x < rem ? x : rem);
    const int c1 = c0 + base + (x < rem ? 1 : 0);
    return XcdRange{c0 + kb, c1, nbx};
}

// R1: cell of every slot and its arrival rank within the cell.  Slots are
// cell-sorted from the previous rebuild, so a wavefront spans few cells:
// one atomic per distinct cell per wavefront.  fused: the particles of each
// chunk of SC_CHUNK cells are counted too (one atomic per distinct chunk per
// wavefront, spread over SC_CHUNK_SPLIT partial counters by block so that
// few wavefronts share an address), so that the scan needs no second kernel.
__global__ __launch_bounds__(SC_TB) void k_bin(const Tile *tiles, const Rep *reps, const int *plan, const int *flag,
                                               int nreps, int ntiles, int fused)
{
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    for (XcdRange it = xcd_range(wk.count); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur));
        const Rep R = ldc(reps, t.rep);
        const int i = t.first + threadIdx.x;
        const bool valid = i < R.n;
        int c = -1;
        if (valid) {
            double4 p = ld4(gp(R.pos) + (i));
            int cc[3];
            if (R.dims == 3)
                cell_of_pos<3>(R, p, cc);
            else
                cell_of_pos<2>(R, p, cc);
            c = cc[0] + R.nc * (cc[1] + R.nc * cc[2]);
            gp(R.cell_of)[i] = c;
        }
        // group the wavefront's lanes by cell (no memory traffic), then every
        // group leader reserves its group's ranks with one atomic, all
        // leaders at once
        const int lane = threadIdx.x % WAVE;
        unsigned long long active = __ballot(valid);
        int leader_of = lane, local = 0, group = 0;
        while (active) {
            const int leader = __ffsll((long long)active) - 1;
            const int lc = __shfl(c, leader, WAVE);
            const unsigned long long same = __ballot(valid && c == lc) & active;
            if ((same >> lane) & 1ull) {
                leader_of = leader;
                local = __popcll(same & ((1ull << lane) - 1ull));
            }
            if (lane == leader) group = __popcll(same);
            active &= ~same;
        }
        int base = 0;
        if (group) base = gatomic_add(&gp(R.cell_cnt)[c], group);
        base = __shfl(base, leader_of, WAVE);
        if (valid) gp(R.rank)[i] = base + local;
        if (fused) {
            const int ch = valid ? c >> SC_CHUNK_SHIFT : -1;
            unsigned long long act = __ballot(valid);
            while (act) {
                const int leader = __ffsll((long long)act) - 1;
                const int lch = __shfl(ch, leader, WAVE);
                const unsigned long long same = __ballot(valid && ch == lch) & act;
                if (lane == leader)
                    gatomic_add(&gp(R.chunk_cnt)[lch * SC_CHUNK_SPLIT + (int)(blockIdx.x % SC_CHUNK_SPLIT)],
                                __popcll(same));
                act &= ~same;
            }
        }
    }
}

// R2a: exclusive scan of the cell counts (SC_CHUNK cells per block, 4 per
// thread); clears the counts for the next rebuild.  fused: the block also
// sums the particle counts of the chunks before its own (counted by k_bin)
// and stores absolute cell starts directly, so R2b and R2c are not
// launched; otherwise it stores chunk-local starts and its chunk's total.
__global__ __launch_bounds__(SC_TB) void k_scan_local(const Rep *reps, const int *plan, const int *flag, int fused)
{
    const int r = rebuild_rep(plan, flag, (int)blockIdx.y);
    if (r < 0) return;
    const Rep R = reps[r];
    const int c0 = blockIdx.x * SC_CHUNK;
    if (c0 >= R.ncell) return;
    const int tid = threadIdx.x;
    int base = 0;
    if (fused) {
        double before = 0.0;  // exact: counts are integers far below 2^53
        for (int q = tid; q < (int)blockIdx.x * SC_CHUNK_SPLIT; q += SC_TB) before += (double)gp(R.chunk_cnt)[q];
        double red[1] = {before};
        block_sum<1>(red);
        base = (int)red[0];
    }
    int v[4];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = c0 + tid * 4 + q;
        v[q] = c < R.ncell ? gp(R.cell_cnt)[c] : 0;
        s += v[q];
    }
    __shared__ int sums[SC_TB];
    sums[tid] = s;
    __syncthreads();
    for (int o = 1; o < SC_TB; o <<= 1) {
        int x = tid >= o ? sums[tid - o] : 0;
        __syncthreads();
        sums[tid] += x;
        __syncthreads();
    }
    int run = base + sums[tid] - s;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = c0 + tid * 4 + q;
        if (c < R.ncell) {
            gp(R.cell_start)[c] = run;
            gp(R.cell_cnt)[c] = 0;
        }
        run += v[q];
    }
    if (!fused && tid == SC_TB - 1) gp(R.chunk_off)[blockIdx.x] = sums[SC_TB - 1];  // chunk total, scanned by R2b
    if (tid == 0 && c0 + SC_CHUNK >= R.ncell) gp(R.cell_start)[R.ncell] = R.n;  // end of the last cell
}

// R2c (after R2b, unfused scans only): chunk offsets added to the
// chunk-local starts, so that every cell start is absolute.
__global__ __launch_bounds__(SC_TB) void k_scan_fold(const Rep *reps, const int *plan, const int *flag)
{
    const int r = rebuild_rep(plan, flag, (int)blockIdx.y);
    if (r < 0) return;
    const Rep R = reps[r];
    const int c0 = blockIdx.x * SC_CHUNK;
    if (c0 >= R.ncell) return;
    const int off = gp(R.chunk_off)[blockIdx.x];
    for (int c = c0 + (int)threadIdx.x; c < min(c0 + SC_CHUNK, R.ncell); c += SC_TB) gp(R.cell_start)[c] += off;
}

// R2b: exclusive scan of the chunk totals (one block per replica).
__global__ __launch_bounds__(SC_TB) void k_scan_top(const Rep *reps, const int *plan, const int *flag)
{
    const int r = rebuild_rep(plan, flag, (int)blockIdx.x);
    if (r < 0) return;
    const Rep R = reps[r];
    const int nch = (R.ncell + SC_CHUNK - 1) / SC_CHUNK;
    __shared__ int sums[SC_TB];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b = 0; b < nch; b += SC_TB) {
        const int q = b + threadIdx.x;
        const int x = q < nch ? gp(R.chunk_off)[q] : 0;
        sums[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < SC_TB; o <<= 1) {
            int y = threadIdx.x >= o ? sums[threadIdx.x - o] : 0;
            __syncthreads();
            sums[threadIdx.x] += y;
            __syncthreads();
        }
        if (q < nch) gp(R.chunk_off)[q] = carry + sums[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 0) carry += sums[SC_TB - 1];
        __syncthreads();
    }
}

// R3: each old slot's cell position; the cell-sorted slot gets the
// particle's original index as its sort key.  fused: the replica's first
// tile clears the chunk counts that R2a has consumed.
__global__ __launch_bounds__(SC_TB) void k_scatter(const Tile *tiles, const Rep *reps, const int *plan, const int *flag,
                                               int nreps, int ntiles, int fused)
{
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    for (XcdRange it = xcd_range(wk.count); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur));
        const Rep R = ldc(reps, t.rep);
        if (fused && t.local == 0) {
            const int nch = (R.ncell + SC_CHUNK - 1) / SC_CHUNK;
            for (int q = threadIdx.x; q < nch * SC_CHUNK_SPLIT; q += SC_TB) gp(R.chunk_cnt)[q] = 0;
        }
        const int i = t.first + threadIdx.x;
        if (i >= R.n) continue;
        gp(R.key)[cstart(R, gp(R.cell_of)[i]) + gp(R.rank)[i]] = gp(R.orig)[i];
    }
}

// R4: final slot = cell start + rank of the original index among the cell's
// members (so the order within a cell is by original index: the slot order
// is a pure function of the positions), and the live state (r/D, v, orig)
// is scattered there.  Inside a step F is dead at this point (recomputed
// next); a rebuild requested between steps (REBUILD_KEEP_F) moves F too,
// through the staging buffer.  disp is zeroed by the build.
__global__ __launch_bounds__(SC_TB) void k_rank_permute(const Tile *tiles, const Rep *reps, const int *plan,
                                                    const int *flag, int nreps, int ntiles)
{
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    for (XcdRange it = xcd_range(wk.count); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur));
        const Rep R = ldc(reps, t.rep);
        const int x = t.first + threadIdx.x, n = R.n;
        if (x >= n) continue;
        const int c = gp(R.cell_of)[x];
        const int s0 = cstart(R, c), e0 = cstart(R, c + 1);
        const int oid = gp(R.orig)[x];
        int rank = 0;
        for (int a = s0; a < e0; a++) rank += gp(R.key)[a] < oid ? 1 : 0;
        const int k = s0 + rank;
        const double4 q = ld4(gp(R.pos) + (x));
        st4(gp(R.pos_alt) + (k), q);
        if (R.bpos)  // the per-slot build's prefilter record of pos_alt[k]
            *(SC_AS1 sc_u32x4v *)(R.bpos + k) =
                sc_u32x4v{fixed_coord(q.x, R.uq), fixed_coord(q.y, R.uq), R.dims == 3 ? fixed_coord(q.z, R.uq) : 0u,
                          __float_as_uint((float)q.w)};
        for (int d = 0; d < R.dims; d++) gp(R.v_alt)[d * n + k] = gp(R.v)[d * n + x];
        if (gp(R.dyn)->rebuild & REBUILD_KEEP_F)
            for (int d = 0; d < R.dims; d++) gp(R.stage)[d * n + k] = gp(R.F)[d * n + x];
        gp(R.orig_alt)[k] = oid;
    }
}

// R5: stencil sweep writing the symmetric short list.  For every slot it
// counts the full-list partners (d2 <= R_v^2, add_pair!'s caller condition)
// for the maxneighbours overflow check and appends each pair that also
// satisfies the half-list predicate d2 < (r_cutoff*sigma_ij + skin)^2
// (src/NeighborLists.jl:92).  It copies the permuted state back and zeroes
// the displacements (src/NeighborLists.jl:35-42).
//
// Cell mode (stencil +-1): one wavefront per run of up to R.cx cells along x
// stages the (cx+2) x 3^(dims-1) window cells' particles in LDS once, then
// every own particle sweeps its 3^dims cells from LDS.  Windows larger than
// SC_WMAX particles read the same candidates from global memory instead.
template <int DIMS>
__device__ inline void build_one(const Rep &R, int k, const double4 p, int &nfull, int &cnt, int j, const double4 q)
{
    double dx, dy, dz;
    double d2 = sqdist<DIMS>(R, p, q, dx, dy, dz);
    if (d2 <= R.Rv2) {
        nfull++;
        double md = (R.pot == 2) ? (p.w + q.w) * 0.5 * (1.0 - R.zeta * fabs(p.w - q.w)) : R.lj_md;
        double rh = R.rcut * md + R.skin;
        if (d2 < rh * rh) {
            if (cnt < R.cap) gp(R.nbr)[(size_t)cnt * R.n + k] = j;
            cnt++;
        }
    }
}

// The permuted state of slot k moved back from the _alt buffers (p: its
// position, pos_alt[k]).
template <int DIMS>
__device__ inline void build_copy_state(const Rep &R, int k, const double4 p)
{
    const int n = R.n;
    st4(gp(R.pos) + (k), p);
    put_shadow(R, k, p);
    const bool keepF = (gp(R.dyn)->rebuild & REBUILD_KEEP_F) != 0;
    for (int c = 0; c < DIMS; c++) {
        gp(R.v)[c * n + k] = gp(R.v_alt)[c * n + k];
        gp(R.disp)[c * n + k] = 0.0;
        if (keepF) gp(R.F)[c * n + k] = gp(R.stage)[c * n + k];
    }
    gp(R.orig)[k] = gp(R.orig_alt)[k];
}

template <int DIMS>
__device__ inline void build_finish_slot(const Rep &R, int k, const double4 p, int nfull, int cnt)
{
    if (nfull > R.cap) gatomic_or(&gp(R.dyn)->err, SC_EBIT_OVERFLOW);
    gp(R.nn)[k] = cnt < R.cap ? cnt : R.cap;
    build_copy_state<DIMS>(R, k, p);
}

// LPP lanes cooperate on one particle: each takes every LPP-th candidate of
// a row range, and the accepted entries are compacted with a ballot in
// candidate order, so the list is identical to a serial sweep's.
struct BuildLds {
    double x[SC_WMAX], y[SC_WMAX], z[SC_WMAX], d[SC_WMAX];  // SoA: conflict-free 8-B reads
    int j[SC_WMAX];
    int coff[WAVE + 1];
    int cglob[WAVE];
};

template <int DIMS, int LPP>
__device__ void build_run(const Rep &R, int run, BuildLds &S)
{
    const int nc = R.nc, CX = R.cx, tid = threadIdx.x;
    const int rpr = (nc + CX - 1) / CX;
    const int rx = run % rpr, rest = run / rpr;
    const int cy = rest % nc, cz = DIMS == 3 ? rest / nc : 0;
    const int x0 = rx * CX, cxr = min(CX, nc - x0);
    const int wx = cxr + 2;
    const int ncw = wx * (DIMS == 3 ? 9 : 3);
    __syncthreads();  // the previous run's readers are done with the LDS
    if (tid < WAVE) {
        int cnt = 0, gs = 0;
        if (tid < ncw) {
            const int w = tid / wx, xi = tid % wx;
            const int oy = (w % 3) - 1, oz = DIMS == 3 ? (w / 3) - 1 : 0;
            const int x = wrapc(x0 - 1 + xi, nc), y = wrapc(cy + oy, nc), z = DIMS == 3 ? wrapc(cz + oz, nc) : 0;
            const int c = x + nc * (y + nc * z);
            gs = cstart(R, c);
            cnt = cstart(R, c + 1) - gs;
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(incl, o, WAVE);
            if (tid >= o) incl += y;
        }
        if (tid < ncw) {
            S.coff[tid] = incl - cnt;
            S.cglob[tid] = gs;
        }
        if (tid == WAVE - 1) S.coff[ncw] = incl;
    }
    __syncthreads();
    const int total = S.coff[ncw];
    const bool staged = total <= SC_WMAX;
    if (staged) {
        for (int a = tid; a < total; a += SC_TB) {
            int lo = 0, hi = ncw - 1;  // last window cell q with S.coff[q] <= a
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (S.coff[mid] <= a)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int j = S.cglob[lo] + (a - S.coff[lo]);
            const double4 q = ld4(gp(R.pos_alt) + (j));
            S.x[a] = q.x;
            S.y[a] = q.y;
            if (DIMS == 3) S.z[a] = q.z;
            S.d[a] = q.w;
            S.j[a] = j;
        }
    }
    __syncthreads();
    const int wc = DIMS == 3 ? 4 : 1;  // window row of the run itself
    const int own0 = S.cglob[wc * wx + 1];
    const int own1 = own0 + (S.coff[wc * wx + cxr + 1] - S.coff[wc * wx + 1]);
    const int grp = tid / LPP, l = tid % LPP;
    const int lane = tid % WAVE;
    const int gshift = (lane / LPP) * LPP;
    const unsigned long long gmask = ((1ull << LPP) - 1ull) << gshift;
    const unsigned long long below = ((1ull << lane) - 1ull) & gmask;
    for (int base = own0; base < own1; base += SC_TB / LPP) {
        const int k = base + grp;
        if (k >= own1) continue;
        const double4 p = ld4(gp(R.pos_alt) + (k));
        int cc[3];
        cell_of_pos<DIMS>(R, p, cc);
        const int xi = cc[0] - x0 + 1;
        const int aself = S.coff[wc * wx + 1] + (k - own0);  // own window index
        int nfull = 0, cnt = 0;
        for (int w = 0; w < (DIMS == 3 ? 9 : 3); w++) {
            const int q0 = w * wx + xi - 1;
            const int a0 = S.coff[q0], a1 = S.coff[q0 + 3];
            for (int a = a0 + l; a - l < a1; a += LPP) {
                bool inF = false, inH = false;
                int j = -1;
                if (a < a1 && a != aself) {
                    double qx, qy, qz = 0.0;
                    if (staged) {
                        qx = S.x[a];
                        qy = S.y[a];
                        if (DIMS == 3) qz = S.z[a];
                    } else {
                        const int cq = a < S.coff[q0 + 1] ? q0 : (a < S.coff[q0 + 2] ? q0 + 1 : q0 + 2);
                        j = S.cglob[cq] + (a - S.coff[cq]);
                        const double4 q = ld4(gp(R.pos_alt) + (j));
                        qx = q.x;
                        qy = q.y;
                        qz = q.z;
                    }
                    const double dx = min_image(p.x - qx, R.L, R.invL);
                    const double dy = min_image(p.y - qy, R.L, R.invL);
                    double d2 = dx * dx + dy * dy;
                    if (DIMS == 3) {
                        const double dz = min_image(p.z - qz, R.L, R.invL);
                        d2 = d2 + dz * dz;
                    }
                    inF = d2 <= R.Rv2;
                    if (inF) {
                        double qd;
                        if (staged) {
                            qd = S.d[a];
                            j = S.j[a];
                        } else {
                            qd = ld4(gp(R.pos_alt) + (j)).w;
                        }
                        const double md = (R.pot == 2) ? (p.w + qd) * 0.5 * (1.0 - R.zeta * fabs(p.w - qd)) : R.lj_md;
                        const double rh = R.rcut * md + R.skin;
                        inH = d2 < rh * rh;
                    }
                }
                const unsigned long long mF = __ballot(inF) & gmask;
                const unsigned long long mH = __ballot(inH) & gmask;
                if (inH) {
                    const int slot = cnt + __popcll(mH & below);
                    if (slot < R.cap) gp(R.nbr)[(size_t)slot * R.n + k] = j;
                }
                nfull += __popcll(mF);
                cnt += __popcll(mH);
            }
        }
        if (l == 0) build_finish_slot<DIMS>(R, k, p, nfull, cnt);
    }
}

__global__ __launch_bounds__(SC_TB) void k_build_cells(const Rep *reps, int nreps, int nruns_total)
{
    __shared__ BuildLds S;
    for (XcdRange it = xcd_range(nruns_total); it.cur < it.end; it.cur += it.step) {
        int r = 0;
        while (r + 1 < nreps && ldc(reps, r + 1).run0 <= it.cur) r++;
        const Rep R = ldc(reps, r);
        if (R.build_kind != BUILD_CELLS || !need_rebuild(R)) continue;
        if (R.dims == 3)
            build_run<3, 8>(R, it.cur - R.run0, S);
        else
            build_run<2, 4>(R, it.cur - R.run0, S);
    }
}

// The FP64 predicates of slot k's pair with slot j (reference minimum image),
// for candidates the prefilter cannot decide: bit 0 full list, bit 1 half list.
template <int DIMS>
__device__ __forceinline__ int exact_pair(const Rep &R, int k, int j)
{
    const double4 p = ld4(gp(R.pos_alt) + (k)), q = ld4(gp(R.pos_alt) + (j));
    double dx, dy, dz;
    const double d2 = sqdist<DIMS>(R, p, q, dx, dy, dz);
    if (!(d2 <= R.Rv2)) return 0;
    const double md = (R.pot == 2) ? (p.w + q.w) * 0.5 * (1.0 - R.zeta * fabs(p.w - q.w)) : R.lj_md;
    const double rh = R.rcut * md + R.skin;
    return d2 < rh * rh ? 3 : 1;
}

// Certain-pass / certain-fail bounds on the squared half-list radius of the
// Berthier pair (sk, sj) in fixed-point units, from the FP32 diameters (the
// evaluation the prefilter margins are set for).
__device__ __forceinline__ void prefilter_half_bounds(const Rep &R, float sk, float sj, float &lo, float &hi)
{
    const float t = fmaf(-R.bf_zeta, fabsf(sk - sj), 1.0f);
    const float rh = fmaf(R.bf_c1, (sk + sj) * t, R.bf_c2);  // (rcut sigma_ij + skin) S
    const float rh2 = rh * rh;
    lo = fmaf(rh2, R.bf_hlo, -R.bf_m);
    hi = fmaf(rh2, R.bf_hhi, R.bf_m);
}

// Fixed-point prefilter (R.bf_on): the pair's separation from the wrapping
// differences of the rebuild-time fixed-point coordinates (the minimum image
// without a shift), squared and summed in FP32, and the half-list radius in
// FP32 from the FP32 diameters; a result within the margins of a threshold
// is decided by exact_pair.  The outcome (bit 0 full list, bit 1 half list)
// is exactly the FP64 predicates'.
template <int DIMS>
__device__ __forceinline__ int prefilter_q(const Rep &R, int k, const uint4 &pk, float sk, int j, const uint4 &q)
{
    const float fx = (float)(int)(pk.x - q.x);
    const float fy = (float)(int)(pk.y - q.y);
    float d2 = fx * fx + fy * fy;
    if (DIMS == 3) {
        const float fz = (float)(int)(pk.z - q.z);
        d2 = d2 + fz * fz;
    }
    if (d2 > R.bf_hi) return 0;
    int res = 1;
    bool exact = d2 >= R.bf_lo;
    if (!exact) {
        float lo = R.bf_rh2lo, hi = R.bf_rh2hi;
        if (R.pot == 2) prefilter_half_bounds(R, sk, __uint_as_float(q.w), lo, hi);
        if (d2 < lo)
            res = 3;
        else
            exact = d2 <= hi;
    }
    if (exact) res = exact_pair<DIMS>(R, k, j);
    return res;
}

template <int DIMS>
__device__ __forceinline__ int prefilter_pair(const Rep &R, int k, const uint4 &pk, float sk, int j)
{
    return prefilter_q<DIMS>(R, k, pk, sk, j, ldq(R.bpos + j));
}

// Cell-image displacement.  With a stencil of +-s cells (cell side w > R_v/s,
// at least 2s+1 cells per side), a candidate in the neighbour cell reached by
// a wrapped offset is displaced by dx_raw - S*L with S in {-1,0,1} the wrap
// of that offset.  For every candidate that can pass d2 <= R_v^2, rint(dx_raw/L)
// == S exactly (|dx_raw - S*L| <= R_v < L/2), so the result is bit-identical to
// dx_raw - rint(dx_raw/L)*L (src/AuxiliaryFunctions.jl:313); candidates for
// which the two images differ are farther than L - (2s+1)w/... > R_v under both
// (|dx_raw - S*L| <= (s+1)w < L - R_v), so no predicate changes.
template <int DIMS>
__device__ inline void build_one_img(const Rep &R, int k, const double4 p, int &nfull, int &cnt, int j,
                                     const double4 q, double sx, double sy, double sz)
{
    const double dx = (p.x - q.x) - sx;
    const double dy = (p.y - q.y) - sy;
    double d2 = dx * dx + dy * dy;
    if (DIMS == 3) {
        const double dz = (p.z - q.z) - sz;
        d2 = d2 + dz * dz;
    }
    if (d2 <= R.Rv2) {
        nfull++;
        const double md = (R.pot == 2) ? (p.w + q.w) * 0.5 * (1.0 - R.zeta * fabs(p.w - q.w)) : R.lj_md;
        const double rh = R.rcut * md + R.skin;
        if (d2 < rh * rh) {
            if (cnt < R.cap) gp(R.nbr)[(size_t)cnt * R.n + k] = j;
            cnt++;
        }
    }
}

// wrapped cell coordinate and the image shift S*L of the wrap
__device__ inline int wrap_shift(int raw, int nc, double L, double &sL)
{
    if (raw < 0) {
        sL = -L;
        return raw + nc;
    }
    if (raw >= nc) {
        sL = L;
        return raw - nc;
    }
    sL = 0.0;
    return raw;
}

// Candidates in the contiguous slot range [b, e) at image shift (sx, sy, sz).
// The loop is deliberately not unrolled: the sweep is latency-bound, and
// the registers an unrolled batch of loads holds cost more (one wavefront
// per SIMD less) than the loads it keeps in flight.
template <int DIMS>
__device__ inline void build_range(const Rep &R, int k, const double4 p, int &nfull, int &cnt, int b, int e,
                                   double sx, double sy, double sz)
{
    for (int j = b; j < e; j++)
        if (j != k) build_one_img<DIMS>(R, k, p, nfull, cnt, j, ld4(gp(R.pos_alt) + (j)), sx, sy, sz);
}

// The stencil of slot k's cell as (2s+1)^(dims-1) rows of cells along x.
// A row's cells are adjacent cell ids, hence one contiguous slot range,
// split where the row wraps the box (each part with its own image shift);
// the parts come in the order of a cell-by-cell sweep.  seg(b, e, sx, sy,
// sz) is called once per part.
// Each row keeps only the cells that reach within R_v of the particle: in
// cell units the particle sits at f in [0, 1] of its cell, so a cell at
// offset o is at least max(0, o - f, f - (o + 1)) away along that axis; rows
// whose y/z gap exceeds R_v are skipped and the x run is cut to
// [floor(f_x - r_x), floor(f_x + r_x)], r_x^2 = R_v^2 - gap^2 (with a 1e-9
// slack far above the rounding of these quantities).  Only candidates
// farther than R_v -- which no predicate accepts -- are dropped, and the
// order of the rest is unchanged.
__device__ inline double cell_gap64(int o, double f)
{
    const double g = o > 0 ? (double)o - f : (o < 0 ? f - (double)(o + 1) : 0.0);
    return g > 1e-9 ? g - 1e-9 : 0.0;
}

template <int DIMS, class Seg>
__device__ inline void for_stencil_rows(const Rep &R, const double4 p, Seg &&seg)
{
    int cc[3];
    cell_of_pos<DIMS>(R, p, cc);
    const int s = R.srange, nc = R.nc;
    const int zr = DIMS == 3 ? s : 0;
    const double fx = p.x * R.cinv - (double)cc[0], fy = p.y * R.cinv - (double)cc[1];
    const double fz = DIMS == 3 ? p.z * R.cinv - (double)cc[2] : 0.0;
    const double rc2 = R.Rv2 * R.cinv * R.cinv * (1.0 + 1e-9) + 1e-9;  // R_v^2 in cell units
    for (int oz = -zr; oz <= zr; oz++) {
        double sz = 0.0;
        const int cz = DIMS == 3 ? wrap_shift(cc[2] + oz, nc, R.L, sz) : 0;
        const double gz = DIMS == 3 ? cell_gap64(oz, fz) : 0.0;
        for (int oy = -s; oy <= s; oy++) {
            const double gy = cell_gap64(oy, fy);
            const double g2 = gy * gy + gz * gz;
            if (g2 > rc2) continue;
            const double rx = sqrt(rc2 - g2) + 1e-9;
            const int lo = max(-s, (int)floor(fx - rx)), hi = min(s, (int)floor(fx + rx));
            if (lo > hi) continue;
            double sy;
            const int cy = wrap_shift(cc[1] + oy, nc, R.L, sy);
            const int row = nc * (cy + nc * cz);
            const int x0 = cc[0] + lo, x1 = cc[0] + hi;
            if (x0 < 0)  // cells seen across the low face
                seg(cstart(R, row + x0 + nc), cstart(R, row + min(x1, -1) + nc + 1), -R.L, sy, sz);
            const int b0 = max(x0, 0), b1 = min(x1, nc - 1);
            if (b0 <= b1) seg(cstart(R, row + b0), cstart(R, row + b1 + 1), 0.0, sy, sz);
            if (x1 >= nc)  // cells seen across the high face
                seg(cstart(R, row + max(x0, nc) - nc), cstart(R, row + x1 - nc + 1), R.L, sy, sz);
        }
    }
}

// One thread per slot (stencil +-srange, or all pairs when srange == 0).
template <int DIMS>
__device__ void build_slot(const Rep &R, int k)
{
    const double4 p = ld4(gp(R.pos_alt) + (k));
    int nfull = 0, cnt = 0;
    if (R.srange == 0) {  // single cell: every other slot, reference minimum image
        for (int j = 0; j < R.n; j++)
            if (j != k) build_one<DIMS>(R, k, p, nfull, cnt, j, ld4(gp(R.pos_alt) + (j)));
    } else {
        for_stencil_rows<DIMS>(R, p, [&](int b, int e, double sx, double sy, double sz) {
            build_range<DIMS>(R, k, p, nfull, cnt, b, e, sx, sy, sz);
        });
    }
    build_finish_slot<DIMS>(R, k, p, nfull, cnt);
}

// The stencil rows of for_stencil_rows as slot ranges only (the prefilter
// needs no image shift), trimmed in FP32 with 1e-4 cell units of slack toward
// inclusion, far above the FP32 rounding of these quantities: a superset of
// the FP64 trim's candidates, all within R_v among the extra ones dropped by
// the predicates, in the same order.  (fx, fy, fz): the slot's fractional
// position in its cell cc.
__device__ inline float cell_gap32(int o, float f)
{
    const float g = o > 0 ? (float)o - f : (o < 0 ? f - (float)(o + 1) : 0.0f);
    return g > 1e-4f ? g - 1e-4f : 0.0f;
}

// Square root for the x extent of a stencil row: the hardware estimate
// (within 1 ulp, below 1e-6 for the arguments here, which are at most
// (s + 1)^2 in cell units) instead of the correctly rounded sequence, which
// costs ~15 more instructions.  Every caller widens the extent by 1e-4 cell
// units, so the cells it lists still cover the radius, and the list entries
// are decided by the exact predicates of the cells' particles alone.
__device__ __forceinline__ float sqrt_extent(float x)
{
    return __builtin_amdgcn_sqrtf(x);
}

template <int DIMS, class Seg>
__device__ __forceinline__ void for_stencil_ranges(const Rep &R, int c0, int c1, int c2, float fx, float fy, float fz, Seg &&seg)
{
    const int s = R.srange, nc = R.nc;
    const int zr = DIMS == 3 ? s : 0;
    const float rc2 = (float)(R.Rv2 * R.cinv * R.cinv) * (1.0f + 1e-4f) + 1e-4f;
    for (int oz = -zr; oz <= zr; oz++) {
        const int cz = DIMS == 3 ? wrapc(c2 + oz, nc) : 0;
        const float gz = DIMS == 3 ? cell_gap32(oz, fz) : 0.0f;
        for (int oy = -s; oy <= s; oy++) {
            const float gy = cell_gap32(oy, fy);
            const float g2 = gy * gy + gz * gz;
            if (g2 > rc2) continue;
            const float rx = sqrt_extent(rc2 - g2) + 1e-4f;
            const int lo = max(-s, (int)floorf(fx - rx)), hi = min(s, (int)floorf(fx + rx));
            if (lo > hi) continue;
            const int row = nc * (wrapc(c1 + oy, nc) + nc * cz);
            const int x0 = c0 + lo, x1 = c0 + hi;
            if (x0 < 0) seg(cstart(R, row + x0 + nc), cstart(R, row + min(x1, -1) + nc + 1));
            const int b0 = max(x0, 0), b1 = min(x1, nc - 1);
            if (b0 <= b1) seg(cstart(R, row + b0), cstart(R, row + b1 + 1));
            if (x1 >= nc) seg(cstart(R, row + max(x0, nc) - nc), cstart(R, row + x1 - nc + 1));
        }
    }
}

// Cell and fractional in-cell position of slot k (the stencil's origin).
template <int DIMS>
__device__ inline void slot_cell(const Rep &R, int k, int &c0, int &c1, int &c2, float &fx, float &fy, float &fz)
{
    const double4 p = ld4(gp(R.pos_alt) + (k));
    c0 = cell_coord(p.x, R);
    c1 = cell_coord(p.y, R);
    c2 = DIMS == 3 ? cell_coord(p.z, R) : 0;
    fx = (float)(p.x * R.cinv - (double)c0);
    fy = (float)(p.y * R.cinv - (double)c1);
    fz = DIMS == 3 ? (float)(p.z * R.cinv - (double)c2) : 0.0f;
}

// One thread per slot with the fixed-point prefilter (R.bf_on, stencil >= 1):
// the same candidates in the same order, so the same list.
template <int DIMS>
__device__ void build_slot_fx(const Rep &R, int k)
{
    int c0, c1, c2;
    float fx, fy, fz;
    slot_cell<DIMS>(R, k, c0, c1, c2, fx, fy, fz);
    const uint4 pk = ldq(R.bpos + k);
    const float sk = __uint_as_float(pk.w);
    int nfull = 0, cnt = 0;
    // SC_BUILD_UNROLL candidate records are loaded together (indices clamped
    // to the range, so every load is in bounds) and then tested in order
    for_stencil_ranges<DIMS>(R, c0, c1, c2, fx, fy, fz, [&](int b, int e) __attribute__((always_inline)) {
        for (int j0 = b; j0 < e; j0 += SC_BUILD_UNROLL) {
            uint4 q[SC_BUILD_UNROLL];
#pragma unroll
            for (int u = 0; u < SC_BUILD_UNROLL; u++) q[u] = ldq(R.bpos + min(j0 + u, e - 1));
#pragma unroll
            for (int u = 0; u < SC_BUILD_UNROLL; u++) {
                const int j = j0 + u;
                if (j >= e || j == k) continue;
                const int res = prefilter_q<DIMS>(R, k, pk, sk, j, q[u]);
                nfull += res & 1;
                if (res & 2) {
                    if (cnt < R.cap) gp(R.nbr)[(size_t)cnt * R.n + k] = j;
                    cnt++;
                }
            }
        }
    });
    build_finish_slot<DIMS>(R, k, ld4(gp(R.pos_alt) + (k)), nfull, cnt);
}

// LPP lanes per slot: lane l of the group takes candidates b+l, b+l+LPP, ...
// of a range; accepted entries are placed with a ballot in candidate order,
// so the list is identical to the one-thread sweep's, with LPP candidate
// loads in flight per slot.  IMG: stencil image shift; else the reference
// minimum image (all-pairs mode).
template <int DIMS, int LPP, bool IMG>
__device__ inline void build_segment(const Rep &R, int k, const double4 p, int b, int e, double sx, double sy,
                                     double sz, int l, unsigned long long gmask, unsigned long long below,
                                     int &nfull, int &cnt)
{
    for (int a0 = b; a0 < e; a0 += LPP) {
        const int a = a0 + l;
        bool inF = false, inH = false;
        if (a < e && a != k) {
            const double4 q = ld4(gp(R.pos_alt) + (a));
            double d2;
            if (IMG) {
                const double dx = (p.x - q.x) - sx;
                const double dy = (p.y - q.y) - sy;
                d2 = dx * dx + dy * dy;
                if (DIMS == 3) {
                    const double dz = (p.z - q.z) - sz;
                    d2 = d2 + dz * dz;
                }
            } else {
                double dx, dy, dz;
                d2 = sqdist<DIMS>(R, p, q, dx, dy, dz);
            }
            inF = d2 <= R.Rv2;
            if (inF) {
                const double md = (R.pot == 2) ? (p.w + q.w) * 0.5 * (1.0 - R.zeta * fabs(p.w - q.w)) : R.lj_md;
                const double rh = R.rcut * md + R.skin;
                inH = d2 < rh * rh;
            }
        }
        const unsigned long long mF = __ballot(inF) & gmask;
        const unsigned long long mH = __ballot(inH) & gmask;
        if (inH) {
            const int slot = cnt + __popcll(mH & below);
            if (slot < R.cap) gp(R.nbr)[(size_t)slot * R.n + k] = a;
        }
        nfull += __popcll(mF);
        cnt += __popcll(mH);
    }
}

// The same with the fixed-point prefilter (R.bf_on).
template <int DIMS, int LPP>
__device__ inline void build_segment_fx(const Rep &R, int k, const uint4 &pk, float sk, int b, int e, int l,
                                        unsigned long long gmask, unsigned long long below, int &nfull, int &cnt)
{
    for (int a0 = b; a0 < e; a0 += LPP) {
        const int a = a0 + l;
        const int res = (a < e && a != k) ? prefilter_pair<DIMS>(R, k, pk, sk, a) : 0;
        const bool inF = (res & 1) != 0, inH = (res & 2) != 0;
        const unsigned long long mF = __ballot(inF) & gmask;
        const unsigned long long mH = __ballot(inH) & gmask;
        if (inH) {
            const int slot = cnt + __popcll(mH & below);
            if (slot < R.cap) gp(R.nbr)[(size_t)slot * R.n + k] = a;
        }
        nfull += __popcll(mF);
        cnt += __popcll(mH);
    }
}

template <int DIMS, int LPP>
__device__ void build_slot_lanes(const Rep &R, int k, int l, unsigned long long gmask, unsigned long long below)
{
    const double4 p = ld4(gp(R.pos_alt) + (k));
    int nfull = 0, cnt = 0;
    if (R.srange == 0) {
        build_segment<DIMS, LPP, false>(R, k, p, 0, R.n, 0.0, 0.0, 0.0, l, gmask, below, nfull, cnt);
    } else {
        for_stencil_rows<DIMS>(R, p, [&](int b, int e, double sx, double sy, double sz) {
            build_segment<DIMS, LPP, true>(R, k, p, b, e, sx, sy, sz, l, gmask, below, nfull, cnt);
        });
    }
    if (l == 0) build_finish_slot<DIMS>(R, k, p, nfull, cnt);
}

template <int DIMS, int LPP>
__device__ void build_slot_lanes_fx(const Rep &R, int k, int l, unsigned long long gmask, unsigned long long below)
{
    int c0, c1, c2;
    float fx, fy, fz;
    slot_cell<DIMS>(R, k, c0, c1, c2, fx, fy, fz);
    const uint4 pk = ldq(R.bpos + k);
    const float sk = __uint_as_float(pk.w);
    int nfull = 0, cnt = 0;
    for_stencil_ranges<DIMS>(R, c0, c1, c2, fx, fy, fz, [&](int b, int e) __attribute__((always_inline)) {
        build_segment_fx<DIMS, LPP>(R, k, pk, sk, b, e, l, gmask, below, nfull, cnt);
    });
    if (l == 0) build_finish_slot<DIMS>(R, k, ld4(gp(R.pos_alt) + (k)), nfull, cnt);
}

// R5 (per-slot form) over the planned tiles.  Launch shape: BS threads per
// block, LPP lanes per slot; a work item is BS/LPP consecutive slots of one
// tile.  The lists do not depend on the launch shape.  The tile and replica
// descriptors are read as constant memory (scalar registers), so the
// one-lane sweep fits 59 VGPRs, seven wavefronts per SIMD, with no spills.
template <int BS, int LPP>
__global__ __launch_bounds__(BS) void k_build(const Tile *tiles, const Rep *reps, const int *plan, const int *flag,
                                              int nreps, int ntiles)
{
    constexpr int SPI = BS / LPP;      // slots per work item
    constexpr int IPT = SC_TB / SPI;   // work items per tile
    static_assert(SPI * IPT == SC_TB && LPP <= WAVE, "launch shape");
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    const int lane = threadIdx.x % WAVE, l = threadIdx.x % LPP;
    const unsigned long long gmask = LPP == WAVE ? ~0ull : ((1ull << LPP) - 1ull) << (lane / LPP * LPP);
    const unsigned long long below = ((1ull << lane) - 1ull) & gmask;
    for (XcdRange it = xcd_range(wk.count * IPT); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur / IPT));
        const Rep R = ldc(reps, t.rep);
        if (R.build_kind != BUILD_SLOT || R.bf_on) continue;
        const int k = t.first + (it.cur % IPT) * SPI + (int)threadIdx.x / LPP;
        if (k >= R.n) continue;
        if (LPP == 1) {
            if (R.dims == 3)
                build_slot<3>(R, k);
            else
                build_slot<2>(R, k);
        } else {
            if (R.dims == 3)
                build_slot_lanes<3, LPP>(R, k, l, gmask, below);
            else
                build_slot_lanes<2, LPP>(R, k, l, gmask, below);
        }
    }
}

// The per-slot form with the fixed-point prefilter, for the replicas with
// R.bf_on (k_build skips them); same launch shapes and work items.
template <int BS, int LPP>
__global__ __launch_bounds__(BS) void k_build_fx(const Tile *tiles, const Rep *reps, const int *plan, const int *flag,
                                                 int nreps, int ntiles)
{
    constexpr int SPI = BS / LPP;
    constexpr int IPT = SC_TB / SPI;
    static_assert(SPI * IPT == SC_TB && LPP <= WAVE, "launch shape");
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    const int lane = threadIdx.x % WAVE, l = threadIdx.x % LPP;
    const unsigned long long gmask = LPP == WAVE ? ~0ull : ((1ull << LPP) - 1ull) << (lane / LPP * LPP);
    const unsigned long long below = ((1ull << lane) - 1ull) & gmask;
    for (XcdRange it = xcd_range(wk.count * IPT); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur / IPT));
        const Rep R = ldc(reps, t.rep);
        if (R.build_kind != BUILD_SLOT || !R.bf_on) continue;
        const int k = t.first + (it.cur % IPT) * SPI + (int)threadIdx.x / LPP;
        if (k >= R.n) continue;
        if (LPP == 1) {
            if (R.dims == 3)
                build_slot_fx<3>(R, k);
            else
                build_slot_fx<2>(R, k);
        } else {
            if (R.dims == 3)
                build_slot_lanes_fx<3, LPP>(R, k, l, gmask, below);
            else
                build_slot_lanes_fx<2, LPP>(R, k, l, gmask, below);
        }
    }
}

// R5, per-slot radius form (launch shape {256, 0}; the replicas with
// R.bf_on).  The lists are entry for entry those of build_slot_fx: the same
// candidate order, the same prefilter and FP64 predicates.  What changes is
// which candidates are visited and how a lane walks them.
//  * Search radius per slot.  Every half-list radius of slot k is at most
//    R_k = rcut * max_j sigma_kj + skin over all diameters D_j in [dmin, dmax]:
//    sigma_kj = (D_k + D_j)/2 (1 - zeta |D_k - D_j|) grows with D_j below
//    D_k, and above it while D_j < 1/(2 zeta), so the maximum is at
//    D_j = min(dmax, max(D_k, 1/(2 zeta))).  The stencil rows are cut to
//    min(R_k, R_v) instead of R_v.  The candidates left out are farther than
//    every half-list radius of the slot, so no accepted entry is lost.
//  * Overflow.  The full-list count (pairs within R_v) is bounded by the
//    particles in the stencil rows cut to R_v, which the cell offsets give
//    without visiting them.  A slot whose bound exceeds the row capacity
//    counts its full-list pairs exactly in one more sweep, so the overflow
//    bit is raised exactly when some full-list row is longer than the cap.
//  * Flattened rows.  Each lane lists the slot ranges of its next stencil
//    rows in LDS (up to PX_TMAX parts) and walks them as one stream with four
//    records in flight, so a wavefront runs as long as its longest stream,
//    not the sum over rows of the longest row.  A row's cell offsets are
//    loaded together, and the next row's while a row is listed.
#ifndef SC_PX_TMAX
#define SC_PX_TMAX 12
#endif
#ifndef SC_PX_U
#define SC_PX_U 4
#endif
constexpr int PX_TMAX = SC_PX_TMAX;  // row parts listed per lane before a sweep
constexpr int PX_U = SC_PX_U;        // candidate records in flight per lane

struct PxLds {
    int2 tab[PX_TMAX][SC_TB];  // part e of thread t at [e][t]
};

// Upper bound on every half-list radius of a particle of diameter Dk.
__device__ inline double px_radius(const Rep &R, double Dk)
{
    if (R.pot != 2) return R.rcut * R.lj_md + R.skin;
    const double dmax = fmax((double)R.dmax, Dk);
    const double Dj = R.zeta > 0.0 ? fmin(dmax, fmax(Dk, 0.5 / R.zeta)) : dmax;
    const double md = (Dk + Dj) * 0.5 * (1.0 - R.zeta * fabs(Dk - Dj));
    // 1e-5 relative: R.dmax is the FP32 rounding of the largest diameter
    return (R.rcut * md + R.skin) * (1.0 + 1e-5);
}

// Cell ids bounding the slot ranges of the raw cells [x0, x1] of the cell
// row starting at id `row`, split where the row wraps the box, in the
// cell-sweep order of for_stencil_ranges: part p is [cstart(id[2p]),
// cstart(id[2p+1])); ids stay -1 where a part is absent.
__device__ __forceinline__ void px_run_ids(int row, int x0, int x1, int nc, int *id)
{
    if (x0 > x1) return;
    if (x0 < 0) {
        id[0] = row + x0 + nc;
        id[1] = row + min(x1, -1) + nc + 1;
        if (x1 >= 0) {
            id[2] = row;
            id[3] = row + x1 + 1;
        }
    } else if (x1 >= nc) {
        if (x0 <= nc - 1) {
            id[0] = row + x0;
            id[1] = row + nc;
        }
        id[2] = row + max(x0, nc) - nc;
        id[3] = row + x1 - nc + 1;
    } else {
        id[0] = row + x0;
        id[1] = row + x1 + 1;
    }
}

// x run [lo, hi] of a stencil row at squared y/z gap g2 within radius^2 r2
// (cell units, with the slack of for_stencil_ranges); lo > hi when empty.
__device__ __forceinline__ void px_run(float r2, float g2, float fx, int s, int &lo, int &hi)
{
    lo = 1;
    hi = 0;
    if (g2 > r2) return;
    const float rx = sqrt_extent(r2 - g2) + 1e-4f;
    lo = max(-s, (int)floorf(fx - rx));
    hi = min(s, (int)floorf(fx + rx));
}

// The decisions of prefilter_q with one branch left (the margins' bands,
// where the FP64 test decides): 0 out, 1 full list only, 3 full and half
// list, 4 decide with exact_pair.  Every quantity is formed exactly as in
// prefilter_q; those prefilter_q skips are computed and not used.
template <int DIMS>
__device__ __forceinline__ int prefilter_sel(const Rep &R, const uint4 &pk, float sk, const uint4 &q)
{
    const float fx = (float)(int)(pk.x - q.x);
    const float fy = (float)(int)(pk.y - q.y);
    float d2 = fx * fx + fy * fy;
    if (DIMS == 3) {
        const float fz = (float)(int)(pk.z - q.z);
        d2 = d2 + fz * fz;
    }
    float lo = R.bf_rh2lo, hi = R.bf_rh2hi;
    if (R.pot == 2) prefilter_half_bounds(R, sk, __uint_as_float(q.w), lo, hi);
    const bool hin = d2 < lo;
    const bool band = d2 >= R.bf_lo || (!hin && d2 <= hi);
    return d2 > R.bf_hi ? 0 : (band ? 4 : (hin ? 3 : 1));
}

// Walks the ne listed parts of this lane as one stream of candidates.
template <int DIMS>
__device__ __forceinline__ void px_sweep(const Rep &R, int k, const uint4 &pk, float sk, int ne, const PxLds &S,
                                         int &cnt)
{
    const int tid = threadIdx.x;
    int ei = 0, j = 0, e = 0;
    if (ne > 0) {
        const int2 t = S.tab[0][tid];
        j = t.x;
        e = t.y;
    }
    SC_AS1 int *out = gp(R.nbr) + (size_t)cnt * R.n + k;  // entry cnt of slot k; advanced by one row per store
    while (__ballot(j < e) != 0ull) {
        int jj[PX_U];
        bool ok[PX_U];
#pragma unroll
        for (int u = 0; u < PX_U; u++) {
            ok[u] = j < e;
            jj[u] = ok[u] ? j : k;
            j++;
            if (j >= e && ei + 1 < ne) {
                ei++;
                const int2 t = S.tab[ei][tid];
                j = t.x;
                e = t.y;
            }
        }
        uint4 q[PX_U];
#pragma unroll
        for (int u = 0; u < PX_U; u++) q[u] = ok[u] ? ldq(R.bpos + jj[u]) : pk;  // no load past a lane's stream
#pragma unroll
        for (int u = 0; u < PX_U; u++) {
            int res = (ok[u] && jj[u] != k) ? prefilter_sel<DIMS>(R, pk, sk, q[u]) : 0;
            if (res & 4) res = exact_pair<DIMS>(R, k, jj[u]);
            const int acc = (res >> 1) & 1;
            if (acc && cnt < R.cap) {
                *out = jj[u];
                out += R.n;
            }
            cnt += acc;
        }
    }
}

template <int DIMS>
__device__ void build_tile_px(const Rep &R, const Tile &t, PxLds &S)
{
    const int tid = threadIdx.x, n = R.n;
    const float cs = (float)(R.cinv * R.cinv);
    const float rv2c = (float)(R.Rv2 * R.cinv * R.cinv) * (1.0f + 1e-4f) + 1e-4f;
    // state copy-back in slot order
    const int k = t.first + tid;
    const bool have = k < n;  // lanes without a slot join the wavefront's ballots only
    if (have) build_copy_state<DIMS>(R, k, ld4(gp(R.pos_alt) + (k)));
    int c0 = 0, c1 = 0, c2 = 0;
    float fx = 0.f, fy = 0.f, fz = 0.f, sk = 0.f, rs2c = 0.f;
    uint4 pk = make_uint4(0u, 0u, 0u, 0u);
    if (have) {
        slot_cell<DIMS>(R, k, c0, c1, c2, fx, fy, fz);
        pk = ldq(R.bpos + k);
        sk = __uint_as_float(pk.w);
        const double rk = fmin(px_radius(R, ld4(gp(R.pos_alt) + (k)).w), sqrt(R.Rv2));
        rs2c = (float)(rk * rk) * cs * (1.0f + 1e-4f) + 1e-4f;
    }
    const int s = R.srange, nc = R.nc, side = 2 * s + 1, nrows = side * (DIMS == 3 ? side : 1);
    // Stencil row r (oz outer, oy inner): the slot-range bounds of its parts
    // as cell offsets, v[0..3] for the run cut to the search radius and
    // v[4..7] for the run cut to R_v (empty parts {0, 0}; same: the two runs
    // coincide and v[4..7] is not loaded).  All of a row's offsets are
    // loaded before any is used, and the next row's are in flight while a
    // row is listed.
    auto fetch_row = [&](int r, int (&v)[8], bool &same) __attribute__((always_inline)) {
        int id[8];
#pragma unroll
        for (int q = 0; q < 8; q++) id[q] = -1;
        same = true;
        if (have) {
            const int oz = DIMS == 3 ? r / side - s : 0, oy = r % side - s;
            const float gz = DIMS == 3 ? cell_gap32(oz, fz) : 0.0f, gy = cell_gap32(oy, fy);
            const float g2 = gy * gy + gz * gz;
            int lov, hiv, los, his;
            px_run(rv2c, g2, fx, s, lov, hiv);
            px_run(rs2c, g2, fx, s, los, his);
            const int row = nc * (wrapc(c1 + oy, nc) + nc * (DIMS == 3 ? wrapc(c2 + oz, nc) : 0));
            px_run_ids(row, c0 + los, c0 + his, nc, id);
            same = lov == los && hiv == his;
            if (!same) px_run_ids(row, c0 + lov, c0 + hiv, nc, id + 4);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = id[q] >= 0 ? cstart(R, id[q]) : 0;
    };
    int bound = -1, cnt = 0, ne = 0;  // bound: particles in the R_v rows, self excluded
    int vc[8];
    bool samec;
    fetch_row(0, vc, samec);
    for (int r = 0; r < nrows; r++) {
        int vn[8];
        bool samen = true;
        if (r + 1 < nrows) fetch_row(r + 1, vn, samen);
        if (have) {
            int listed = 0;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int b = vc[2 * p], e = vc[2 * p + 1];
                if (b < e) {
                    S.tab[ne][tid] = make_int2(b, e);
                    ne++;
                    listed += e - b;
                }
            }
            bound += samec ? listed : (vc[5] - vc[4]) + (vc[7] - vc[6]);
        }
        if (r == nrows - 1 || __ballot(ne > PX_TMAX - 3) != 0ull) {
            px_sweep<DIMS>(R, have ? k : t.first, pk, sk, ne, S, cnt);
            ne = 0;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) vc[q] = vn[q];
        samec = samen;
    }
    if (!have) return;
    if (bound > R.cap) {
        // the bound allows an overflow: count the full-list pairs exactly
        int nfull = 0;
        for_stencil_ranges<DIMS>(R, c0, c1, c2, fx, fy, fz, [&](int b, int e) __attribute__((always_inline)) {
            for (int j = b; j < e; j++)
                if (j != k) nfull += prefilter_pair<DIMS>(R, k, pk, sk, j) & 1;
        });
        if (nfull > R.cap) gatomic_or(&gp(R.dyn)->err, SC_EBIT_OVERFLOW);
    }
    gp(R.nn)[k] = cnt < R.cap ? cnt : R.cap;
}

// Held to six wavefronts per SIMD (80 VGPRs, no spills): the store pointer
// of px_sweep would otherwise take it to 81 VGPRs and five wavefronts.
__global__ __launch_bounds__(SC_TB) __attribute__((amdgpu_waves_per_eu(6))) void k_build_px(const Tile *tiles, const Rep *reps, const int *plan,
                                                     const int *flag, int nreps, int ntiles)
{
    __shared__ PxLds S;
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    for (XcdRange it = xcd_range(wk.count); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur));
        const Rep R = ldc(reps, t.rep);
        if (R.build_kind != BUILD_SLOT || !R.bf_on) continue;
        if (R.dims == 3)
            build_tile_px<3>(R, t, S);
        else
            build_tile_px<2>(R, t, S);
        __syncthreads();  // the LDS tables are reused by the next tile
    }
}

// R5 (staged-row form, stencil +-s, s >= 1).  One wave (a block of 64
// threads) per 64 consecutive slots.  Its lanes group by own cell row (y, z)
// -- cell-sorted slots span one row, sometimes two -- and for each of the
// group's (2s+1)^(dims-1) stencil rows the wave stages the candidates of the
// whole raw x range [xmin - s, xmax + s] into LDS with coalesced loads (one
// contiguous slot range per wrap part), with the staged offset of every raw
// cell beside them.  Each lane then sweeps its own x window (cut to the list
// sphere as in for_stencil_rows) from LDS.  Candidates keep the per-slot
// sweep's order (rows in order, raw x ascending, slots ascending) and the
// same exact image-shift d^2 (build_one_img), so the lists are entry for
// entry those of the per-slot sweep; only where the candidates come from
// changes.  Rows longer than RB_CAP candidates are staged in chunks.
constexpr int RB_CAP = 256;   // staged candidates per chunk
constexpr int RB_XMAX = 192;  // raw cells whose offsets fit the LDS table

struct RowsLds {
    double4 q[RB_CAP];
    int off[RB_XMAX + 1];
};

__device__ inline int wave_min_int(int x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, WAVE));
    return x;
}

__device__ inline int wave_max_int(int x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, WAVE));
    return x;
}

// Staged offset of raw cell x (rlo <= x <= rhi + 1) of stencil row `row`: the
// raw range is laid out as wrap part A (x < 0, cells x + nc), part B
// (0 <= x < nc) and part C (x >= nc, cells x - nc), each one slot range.
struct RowParts {
    int sA, nA, sB, nB, sC, nC;
};

__device__ inline int staged_offset(const Rep &R, int row, int x, int rhi, const RowParts &P)
{
    if (x > rhi) return P.nA + P.nB + P.nC;
    if (x < 0) return cstart(R, row + x + R.nc) - P.sA;
    if (x < R.nc) return P.nA + cstart(R, row + x) - P.sB;
    return P.nA + P.nB + cstart(R, row + x - R.nc) - P.sC;
}

template <int DIMS>
__device__ void rows_group(const Rep &R, int k0, RowsLds &S)
{
    const int lane = threadIdx.x, n = R.n, nc = R.nc, s = R.srange;
    const int k = k0 + lane;
    const bool valid = k < n;
    const double4 p = valid ? ld4(gp(R.pos_alt) + (k)) : make_double4(0.0, 0.0, 0.0, 0.0);
    int cc[3];
    cell_of_pos<DIMS>(R, p, cc);
    const double fx = p.x * R.cinv - (double)cc[0], fy = p.y * R.cinv - (double)cc[1];
    const double fz = DIMS == 3 ? p.z * R.cinv - (double)cc[2] : 0.0;
    const double rc2 = R.Rv2 * R.cinv * R.cinv * (1.0 + 1e-9) + 1e-9;  // R_v^2 in cell units
    const int zr = DIMS == 3 ? s : 0;
    int nfull = 0, cnt = 0;
    unsigned long long todo = __ballot(valid);
    while (todo) {  // the distinct own rows of the group (usually one)
        // the row loop's values are wave-uniform: readfirstlane keeps them in
        // scalar registers (and the row loop's branches scalar)
        const int leader = __ffsll((long long)todo) - 1;
        const int ry = __builtin_amdgcn_readfirstlane(__shfl(cc[1], leader, WAVE));
        const int rz = __builtin_amdgcn_readfirstlane(__shfl(cc[2], leader, WAVE));
        const bool mine = valid && cc[1] == ry && cc[2] == rz;
        todo &= ~__ballot(mine);
        const int rlo = __builtin_amdgcn_readfirstlane(wave_min_int(mine ? cc[0] : 1 << 30)) - s;
        const int rhi = __builtin_amdgcn_readfirstlane(wave_max_int(mine ? cc[0] : -(1 << 30))) + s;
        const int ncells = rhi - rlo + 2;  // raw cells plus the end
        for (int oz = -zr; oz <= zr; oz++) {
            double sz = 0.0;
            const int cz = DIMS == 3 ? wrap_shift(rz + oz, nc, R.L, sz) : 0;
            const double gz = DIMS == 3 ? cell_gap64(oz, fz) : 0.0;
            for (int oy = -s; oy <= s; oy++) {
                double sy;
                const int cy = wrap_shift(ry + oy, nc, R.L, sy);
                const int row = nc * (cy + nc * cz);
                RowParts P;
                {
                    const int a1 = min(rhi, -1), b0 = max(rlo, 0), b1 = min(rhi, nc - 1), c0 = max(rlo, nc);
                    P.sA = rlo < 0 ? cstart(R, row + rlo + nc) : 0;
                    P.nA = rlo < 0 ? cstart(R, row + a1 + nc + 1) - P.sA : 0;
                    P.sB = b0 <= b1 ? cstart(R, row + b0) : 0;
                    P.nB = b0 <= b1 ? cstart(R, row + b1 + 1) - P.sB : 0;
                    P.sC = rhi >= nc ? cstart(R, row + c0 - nc) : 0;
                    P.nC = rhi >= nc ? cstart(R, row + rhi - nc + 1) - P.sC : 0;
                }
                const int T = P.nA + P.nB + P.nC;
                // this lane's x window, cut to the list sphere, in raw cells
                int lo = 1, hi = 0;
                if (mine) {
                    const double gy = cell_gap64(oy, fy);
                    const double g2 = gy * gy + gz * gz;
                    if (g2 <= rc2) {
                        const double rx = sqrt(rc2 - g2) + 1e-9;
                        lo = cc[0] + max(-s, (int)floor(fx - rx));
                        hi = cc[0] + min(s, (int)floor(fx + rx));
                    }
                }
                int a0 = 0, a1 = 0;
                const bool tab = ncells <= RB_XMAX + 1;
                __syncthreads();  // the previous row's readers are done with the LDS
                if (tab)
                    for (int x = lane; x < ncells; x += WAVE) S.off[x] = staged_offset(R, row, rlo + x, rhi, P);
                __syncthreads();
                if (lo <= hi) {
                    a0 = tab ? S.off[lo - rlo] : staged_offset(R, row, lo, rhi, P);
                    a1 = tab ? S.off[hi + 1 - rlo] : staged_offset(R, row, hi + 1, rhi, P);
                }
                for (int c0 = 0; c0 < T; c0 += RB_CAP) {
                    const int cn = min(RB_CAP, T - c0);
                    if (c0 > 0) __syncthreads();
                    for (int a = lane; a < cn; a += WAVE) {
                        const int g = c0 + a;
                        const int j = g < P.nA ? P.sA + g : (g < P.nA + P.nB ? P.sB + (g - P.nA) : P.sC + (g - P.nA - P.nB));
                        S.q[a] = ld4(gp(R.pos_alt) + (j));
                    }
                    __syncthreads();
                    const int b0 = max(a0, c0), b1 = min(a1, c0 + cn);
#pragma unroll 1
                    for (int g = b0; g < b1; g++) {
                        int j;
                        double sx;
                        if (g < P.nA) {
                            j = P.sA + g;
                            sx = -R.L;
                        } else if (g < P.nA + P.nB) {
                            j = P.sB + (g - P.nA);
                            sx = 0.0;
                        } else {
                            j = P.sC + (g - P.nA - P.nB);
                            sx = R.L;
                        }
                        if (j != k) build_one_img<DIMS>(R, k, p, nfull, cnt, j, S.q[g - c0], sx, sy, sz);
                    }
                }
            }
        }
    }
    if (valid) build_finish_slot<DIMS>(R, k, p, nfull, cnt);
}

// One kernel per dimensionality (each holds only its own sweep's registers).
template <int DIMS>
__global__ __launch_bounds__(WAVE) void k_build_rows(const Tile *tiles, const Rep *reps, const int *plan,
                                                     const int *flag, int nreps, int ntiles)
{
    __shared__ RowsLds S;
    constexpr int GPT = SC_TB / WAVE;  // 64-slot groups per tile
    const RebuildWork wk = rebuild_work(plan, nreps, flag, ntiles);
    for (XcdRange it = xcd_range(wk.count * GPT); it.cur < it.end; it.cur += it.step) {
        const Tile t = ldc(tiles, wk.tile(it.cur / GPT));
        const Rep R = ldc(reps, t.rep);
        const int k0 = t.first + (it.cur % GPT) * WAVE;
        // uniform over the block (one wave)
        if (R.build_kind != BUILD_ROWS || R.dims != DIMS || k0 >= R.n) continue;
        rows_group<DIMS>(R, k0, S);
    }
}

// R5 (windowed form, any stencil +-s, s <= 3).  A run is cx x cy x cz own
// cells (cz = 1 in 2D); its window -- the own cells and s more cells on
// every side, wrapped -- is staged in LDS once: window-relative FP32
// coordinates with the cell's image shift applied, FP32 D and the slot index.
// Each own particle is swept by WIN_LW lanes: they cut every stencil row
// (rows oz, oy in the per-slot sweep's order) to the run of cells that lie
// within the particle's largest half-list radius, lay the surviving rows end
// to end, and walk that flat candidate list WIN_LW entries at a time.  The
// list order is the per-slot sweep's (rows in order, cells along x in offset
// order, slots in order), minus candidates no predicate could accept, so
// the lists are identical to that sweep's.  The FP32 d^2 decides every
// candidate outside a thin band around R_v^2 and around the half-list radius
// (margins from the coordinate magnitudes, Rep::wlo/whi/hmarg); inside a band
// the reference's FP64 d^2 decides, from global memory.  A window holding
// more than SC_WIN_MAX particles sweeps its own particles from global memory.
constexpr int WIN_LW = 8;                 // lanes per own particle
constexpr int WIN_GROUPS = SC_TB / WIN_LW;
constexpr int WIN_RMAX = 49;              // stencil rows per particle, (2s+1)^2 with s <= 3
constexpr int WIN_RPL = (WIN_RMAX + WIN_LW - 1) / WIN_LW;  // rows per lane
constexpr int WIN_SB = 2;                 // staging loads in flight per thread
constexpr int WIN_ABITS = 11;             // LDS index bits of a row entry (SC_WIN_MAX = 2^11)
static_assert(SC_WIN_MAX == 1 << WIN_ABITS, "row entry packing");

struct WinLds {
    float4 q[SC_WIN_MAX];
    int j[SC_WIN_MAX];
    int coff[SC_WIN_CELLS + 1];
    int cglob[SC_WIN_CELLS];
    int orow[SC_WIN_ROWS + 1];
    int wsum[NWAVES];
    int rows[WIN_GROUPS][WIN_RMAX + 1];  // per lane group: first LDS index | flat offset << WIN_ABITS
};

// Cell id and image shift of window cell w.
template <int DIMS>
__device__ inline int win_cell(const Rep &R, int w, int WX, int WY, int X0, int Y0, int Z0, double &sx, double &sy,
                               double &sz)
{
    const int s = R.srange, nc = R.nc;
    const int wx = w % WX, wy = (w / WX) % WY, wz = w / (WX * WY);
    const int cx = wrap_shift(X0 - s + wx, nc, R.L, sx);
    const int cy = wrap_shift(Y0 - s + wy, nc, R.L, sy);
    int cz = 0;
    sz = 0.0;
    if (DIMS == 3) cz = wrap_shift(Z0 - s + wz, nc, R.L, sz);
    return cx + nc * (cy + nc * cz);
}

// Lower bound on the distance along one axis from window-relative coordinate
// x to window cell c (which spans [c w, (c+1) w)), less a slack eps.
__device__ inline float cell_gap(int c, float x, float w, float eps)
{
    const float g = fmaxf(fmaxf((float)c * w - x, x - (float)(c + 1) * w), 0.0f) - eps;
    return fmaxf(g, 0.0f);
}

// The reference predicates in FP64 (add_pair! and its caller's cutoff).
template <int DIMS>
__device__ inline void win_exact(const Rep &R, const double4 p, int j, bool &inF, bool &inH)
{
    const double4 q = ld4(gp(R.pos_alt) + (j));
    double dx, dy, dz;
    const double d2 = sqdist<DIMS>(R, p, q, dx, dy, dz);
    inF = d2 <= R.Rv2;
    inH = false;
    if (inF) {
        const double md = (R.pot == 2) ? (p.w + q.w) * 0.5 * (1.0 - R.zeta * fabs(p.w - q.w)) : R.lj_md;
        const double rh = R.rcut * md + R.skin;
        inH = d2 < rh * rh;
    }
}

// Window x-cell range [lo, hi] of stencil row (dy, dz) within radius^2 R2 of
// the particle (pf, window cell x wxk); lo > hi when the row lies beyond it.
struct WinGeo {
    float wf, winv, eps;
    int s;
};

__device__ inline void win_row_range(const WinGeo &G, float4 pf, int wxk, float g2, float R2, int &lo, int &hi)
{
    lo = 1;
    hi = 0;
    if (g2 > R2) return;
    const float rx = sqrtf(R2 - g2) + G.eps;
    lo = max(wxk - G.s, (int)floorf((pf.x - rx) * G.winv));
    hi = min(wxk + G.s, (int)floorf((pf.x + rx) * G.winv));
}

template <int DIMS>
__device__ void win_run(const Rep &R, int run, WinLds &S)
{
    static_assert(SC_WIN_CELLS <= 4 * SC_TB, "four window cells per thread");
    constexpr int LW = WIN_LW;
    const int nc = R.nc, s = R.srange, tid = threadIdx.x;
    const int nrx = (nc + R.cx - 1) / R.cx, nry = (nc + R.cy - 1) / R.cy;
    const int rx = run % nrx, ry = (run / nrx) % nry, rz = run / (nrx * nry);
    const int X0 = rx * R.cx, Y0 = ry * R.cy, Z0 = DIMS == 3 ? rz * R.cz : 0;
    const int cxr = min(R.cx, nc - X0), cyr = min(R.cy, nc - Y0), czr = DIMS == 3 ? min(R.cz, nc - Z0) : 1;
    const int WX = cxr + 2 * s, WY = cyr + 2 * s, WZ = DIMS == 3 ? czr + 2 * s : 1;
    const int ncw = WX * WY * WZ;
    __syncthreads();  // the previous run's readers are done with the LDS
    // window cell populations and their exclusive scan, four cells per thread
    int cnt[4], tsum = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int w = tid * 4 + q;
        cnt[q] = 0;
        if (w < ncw) {
            double sx, sy, sz;
            const int c = win_cell<DIMS>(R, w, WX, WY, X0, Y0, Z0, sx, sy, sz);
            const int g0 = cstart(R, c);
            S.cglob[w] = g0;
            cnt[q] = cstart(R, c + 1) - g0;
        }
        tsum += cnt[q];
    }
    const int lane = tid % WAVE, wv = tid / WAVE;
    int incl = tsum;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += y;
    }
    if (lane == WAVE - 1) S.wsum[wv] = incl;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int q = 0; q < NWAVES; q++) {
        const int x = S.wsum[q];
        before += q < wv ? x : 0;
        total += x;
    }
    int off = before + incl - tsum;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int w = tid * 4 + q;
        if (w < ncw) S.coff[w] = off;
        off += cnt[q];
    }
    if (tid == 0) S.coff[ncw] = total;
    __syncthreads();
    // own cell rows (cells X0 .. X0+cxr-1 of one y, z): particle offsets
    const int nrows = cyr * czr;
    if (tid < WAVE) {
        int oc = 0;
        if (tid < nrows) {
            const int wy = s + tid % cyr, wz = DIMS == 3 ? s + tid / cyr : 0;
            const int w0 = (wz * WY + wy) * WX + s;
            oc = S.coff[w0 + cxr] - S.coff[w0];
        }
        int inc = oc;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(inc, o, WAVE);
            if (tid >= o) inc += y;
        }
        if (tid < nrows) S.orow[tid] = inc - oc;
        if (tid == WAVE - 1) S.orow[nrows] = inc;
    }
    // staging: up to SC_WIN_MAX / SC_TB particles per thread, loads issued
    // WIN_SB at a time
    const bool staged = total <= SC_WIN_MAX;
    if (staged) {
        const double ox = (double)(X0 - s) * R.wcell, oy = (double)(Y0 - s) * R.wcell;
        const double oz = DIMS == 3 ? (double)(Z0 - s) * R.wcell : 0.0;
        for (int a0 = tid; a0 < total; a0 += WIN_SB * SC_TB) {
            int wc[WIN_SB], jj[WIN_SB];
            double4 P[WIN_SB];
#pragma unroll
            for (int u = 0; u < WIN_SB; u++) {
                const int a = a0 + u * SC_TB;
                wc[u] = -1;
                jj[u] = 0;
                if (a < total) {
                    int lo = 0, hi = ncw - 1;  // the window cell holding a: last q with coff[q] <= a
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (S.coff[mid] <= a)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    wc[u] = lo;
                    jj[u] = S.cglob[lo] + (a - S.coff[lo]);
                }
            }
#pragma unroll
            for (int u = 0; u < WIN_SB; u++)
                if (wc[u] >= 0) P[u] = ld4(gp(R.pos_alt) + (jj[u]));
#pragma unroll
            for (int u = 0; u < WIN_SB; u++) {
                if (wc[u] < 0) continue;
                const int a = a0 + u * SC_TB;
                double sx, sy, sz;
                (void)win_cell<DIMS>(R, wc[u], WX, WY, X0, Y0, Z0, sx, sy, sz);
                S.q[a] = make_float4((float)((P[u].x + sx) - ox), (float)((P[u].y + sy) - oy),
                                     DIMS == 3 ? (float)((P[u].z + sz) - oz) : 0.0f, (float)P[u].w);
                S.j[a] = jj[u];
            }
        }
    }
    __syncthreads();
    const int nown = S.orow[nrows];
    const int g = tid / LW, l = tid % LW;
    const unsigned long long gmask = ((1ull << LW) - 1ull) << (lane / LW * LW);
    const unsigned long long below = ((1ull << lane) - 1ull) & gmask;
    const float zeta = (float)R.zeta, rcut = (float)R.rcut, skin = (float)R.skin, ljmd = (float)R.lj_md;
    const float wlo = R.wlo, whi = R.whi, hm = R.hmarg;
    WinGeo G;
    G.wf = (float)R.wcell;
    G.winv = (float)(1.0 / R.wcell);
    G.eps = 1e-5f * G.wf + 1e-6f;  // > FP32 coordinate error + cell-boundary rounding
    G.s = s;
    const float rvm = (float)sqrt(R.Rv2) * 1.00001f + G.eps;
    const float R2v = rvm * rvm;
    const float dmax = R.dmax * 1.000001f;
    const int n = R.n, cap = R.cap;
    const int side = 2 * s + 1, nst = DIMS == 3 ? side * side : side;  // stencil rows
    int *rows = S.rows[g];
    for (int base = 0; base < nown; base += WIN_GROUPS) {  // same trip count for every lane of the block
        const int o = base + g;
        const bool active = o < nown;
        int k = 0, aself = 0, wxk = 0, wy0 = 0, wz0 = 0, T = 0, bound = 0;
        double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
        float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
        if (active) {
            int r = 0;
            while (r + 1 < nrows && S.orow[r + 1] <= o) r++;
            wy0 = s + r % cyr;
            wz0 = DIMS == 3 ? s + r / cyr : 0;
            const int w0 = (wz0 * WY + wy0) * WX + s;
            k = S.cglob[w0] + (o - S.orow[r]);
            aself = S.coff[w0] + (o - S.orow[r]);
            if (staged) {
                p = ld4(gp(R.pos_alt) + (k));
                int cc[3];
                cell_of_pos<DIMS>(R, p, cc);
                wxk = cc[0] - X0 + s;
                pf = S.q[aself];
                // largest half-list radius of this particle over every
                // partner diameter in (0, dmax]: sigma_ij rises with D_j up
                // to D_j = D_i, and beyond it up to 1/(2 zeta)
                float smax = ljmd;
                if (R.pot == 2) {
                    const float dst = zeta > 0.0f ? fminf(fmaxf(0.5f / zeta, pf.w), dmax) : dmax;
                    smax = fmaxf(pf.w, (pf.w + dst) * 0.5f * (1.0f - zeta * fabsf(pf.w - dst)));
                }
                const float rhm = fminf((rcut * smax + skin) * 1.00001f + G.eps, rvm);
                const float R2h = rhm * rhm;
                // this lane's block of stencil rows: candidate counts within
                // the half-list radius (and LDS starts) and the R_v bound
                const int B = (nst + LW - 1) / LW;  // rows per lane, lane l takes [l B, l B + B)
                int ra[WIN_RPL], rn[WIN_RPL], mine = 0;
#pragma unroll
                for (int q = 0; q < WIN_RPL; q++) {
                    const int rr = l * B + q;
                    ra[q] = 0;
                    rn[q] = 0;
                    if (q < B && rr < nst) {
                        const int dy = rr % side - s, dz = DIMS == 3 ? rr / side - s : 0;
                        const float gy = cell_gap(wy0 + dy, pf.y, G.wf, G.eps);
                        const float gz = DIMS == 3 ? cell_gap(wz0 + dz, pf.z, G.wf, G.eps) : 0.0f;
                        const float g2 = gy * gy + gz * gz;
                        const int rowc = ((wz0 + dz) * WY + (wy0 + dy)) * WX;
                        int lo, hi;
                        win_row_range(G, pf, wxk, g2, R2v, lo, hi);
                        if (lo <= hi) bound += S.coff[rowc + hi + 1] - S.coff[rowc + lo];
                        win_row_range(G, pf, wxk, g2, R2h, lo, hi);
                        if (lo <= hi) {
                            const int a0 = S.coff[rowc + lo], na = S.coff[rowc + hi + 1] - a0;
                            if (na > 0) {
                                ra[q] = a0;
                                rn[q] = na;
                            }
                        }
                    }
                    mine += rn[q];
                }
                // exclusive scan of the lanes' counts within the group
                int inc = mine;
#pragma unroll
                for (int d = 1; d < LW; d <<= 1) {
                    const int y = __shfl_up(inc, d, LW);
                    if (l >= d) inc += y;
                }
                T = __shfl(inc, LW - 1, LW);
                int P = inc - mine;
#pragma unroll
                for (int q = 0; q < WIN_RPL; q++) {
                    const int rr = l * B + q;
                    if (q < B && rr < nst) rows[rr] = ra[q] | (P << WIN_ABITS);
                    P += rn[q];
                }
                if (l == LW - 1) rows[nst] = T << WIN_ABITS;
#pragma unroll
                for (int d = 1; d < LW; d <<= 1) bound += __shfl_xor(bound, d, LW);
            }
        }
        __syncthreads();
        if (active) {
            if (!staged) {
                build_slot_lanes<DIMS, LW>(R, k, l, gmask, below);
            } else {
                int cnt_h = 0, r = 0;
                for (int c0 = 0; c0 < T; c0 += LW) {
                    const int c = c0 + l;
                    bool inH = false;
                    int a = -1;
                    if (c < T) {
                        while ((rows[r + 1] >> WIN_ABITS) <= c) r++;
                        const int e = rows[r];
                        a = (e & (SC_WIN_MAX - 1)) + (c - (e >> WIN_ABITS));
                        if (a != aself) {
                            const float4 q = S.q[a];
                            const float ex = pf.x - q.x, ey = pf.y - q.y;
                            float d2 = ex * ex + ey * ey;
                            if (DIMS == 3) {
                                const float ez = pf.z - q.z;
                                d2 = d2 + ez * ez;
                            }
                            if (d2 <= whi) {
                                const float md = (R.pot == 2) ? (pf.w + q.w) * 0.5f * (1.0f - zeta * fabsf(pf.w - q.w))
                                                              : ljmd;
                                const float rh = rcut * md + skin;
                                const float rh2 = rh * rh;
                                if (d2 < rh2 - hm && d2 <= wlo) {
                                    inH = true;
                                } else if (d2 < rh2 + hm) {
                                    bool inF;
                                    win_exact<DIMS>(R, p, S.j[a], inF, inH);
                                }
                            }
                        }
                    }
                    const unsigned long long mH = __ballot(inH) & gmask;
                    if (inH) {
                        const int slot = cnt_h + __popcll(mH & below);
                        if (slot < cap) gp(R.nbr)[(size_t)slot * n + k] = S.j[a];
                    }
                    cnt_h += __popcll(mH);
                }
                // The full-list length matters only against maxneighbours:
                // the candidate count within R_v bounds it; count exactly only
                // when the bound exceeds the capacity.
                int nfull = bound;
                if (bound > cap) {
                    nfull = 0;
                    for (int rr = 0; rr < nst; rr++) {
                        const int dy = rr % side - s, dz = DIMS == 3 ? rr / side - s : 0;
                        const float gy = cell_gap(wy0 + dy, pf.y, G.wf, G.eps);
                        const float gz = DIMS == 3 ? cell_gap(wz0 + dz, pf.z, G.wf, G.eps) : 0.0f;
                        int lo, hi;
                        win_row_range(G, pf, wxk, gy * gy + gz * gz, R2v, lo, hi);
                        if (lo > hi) continue;
                        const int rowc = ((wz0 + dz) * WY + (wy0 + dy)) * WX;
                        const int a1 = S.coff[rowc + hi + 1];
                        for (int b = S.coff[rowc + lo]; b < a1; b += LW) {
                            const int a = b + l;
                            bool inF = false;
                            if (a < a1 && a != aself) {
                                const float4 q = S.q[a];
                                const float ex = pf.x - q.x, ey = pf.y - q.y;
                                float d2 = ex * ex + ey * ey;
                                if (DIMS == 3) {
                                    const float ez = pf.z - q.z;
                                    d2 = d2 + ez * ez;
                                }
                                if (d2 <= wlo) {
                                    inF = true;
                                } else if (d2 <= whi) {
                                    bool inH2;
                                    win_exact<DIMS>(R, p, S.j[a], inF, inH2);
                                }
                            }
                            nfull += __popcll(__ballot(inF) & gmask);
                        }
                    }
                }
                if (l == 0) build_finish_slot<DIMS>(R, k, p, nfull, cnt_h);
            }
        }
        __syncthreads();  // the group's row table is rewritten by the next own particle
    }
}

// R5 (windowed form) over the planned (replica, run) pairs, or, with no plan
// (a single replica), all of the replica's runs when its flag is set.
__global__ __launch_bounds__(SC_TB, 3) void k_build_win(const Rep *reps, const int *plan, const int *flag, int nreps,
                                                        int ntiles)
{
    __shared__ WinLds S;
    const int *rl = plan ? plan + 4 + nreps + ntiles : nullptr;
    const int count = plan ? plan[2]
                           : ((*gp(flag) != 0 && reps[0].build_kind == BUILD_WIN) ? reps[0].nruns : 0);
    for (XcdRange it = xcd_range(count); it.cur < it.end; it.cur += it.step) {
        const int r = rl ? cp(rl)[2 * it.cur] : 0;
        const int run = rl ? cp(rl)[2 * it.cur + 1] : it.cur;
        const Rep R = ldc(reps, r);
        if (R.dims == 3)
            win_run<3>(R, run, S);
        else
            win_run<2>(R, run, S);
    }
}

// Full rows as the last rebuild stored them (src/NeighborLists.jl:82-104:
// the reference fills its full list at rebuild time and leaves it until the
// next rebuild), for the reference-layout list export.  pos_alt is written
// only by the rebuild's permutation, so between rebuilds it holds the
// rebuild-time positions in the current slot order; each slot's cell is the
// cell it was binned into.
template <int DIMS>
__device__ void full_rows_tile(const Rep &R, const Tile &t, int *rows, int *counts, int cap)
{
    const int n = R.n, k = t.first + threadIdx.x;
    if (k >= n) return;
    const double4 p = ld4(gp(R.pos_alt) + (k));
    int cc[3];
    cell_of_pos<DIMS>(R, p, cc);
    const int s = R.srange;
    const int zr = DIMS == 3 ? s : 0;
    int cnt = 0;
    for (int oz = -zr; oz <= zr; oz++)
        for (int oy = -s; oy <= s; oy++)
            for (int ox = -s; ox <= s; ox++) {
                const int c = wrapc(cc[0] + ox, R.nc) +
                              R.nc * (wrapc(cc[1] + oy, R.nc) + R.nc * (DIMS == 3 ? wrapc(cc[2] + oz, R.nc) : 0));
                for (int j = cstart(R, c), je = cstart(R, c + 1); j < je; j++) {
                    if (j == k) continue;
                    double dx, dy, dz;
                    double d2 = sqdist<DIMS>(R, p, ld4(gp(R.pos_alt) + (j)), dx, dy, dz);
                    if (d2 <= R.Rv2) {
                        if (cnt < cap) rows[(size_t)gp(R.orig)[k] * cap + cnt] = gp(R.orig)[j];
                        cnt++;
                    }
                }
            }
    counts[gp(R.orig)[k]] = cnt;
}

__global__ __launch_bounds__(SC_TB) void k_full_rows(const Tile *tiles, const Rep *reps, int *rows, int *counts, int cap)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (R.dims == 3)
        full_rows_tile<3>(R, t, rows, counts, cap);
    else
        full_rows_tile<2>(R, t, rows, counts, cap);
}

// Reference-layout rows built on the device (the allocation-free list
// export): for the slot of particle a, row a of the column-major
// maxneigh x N Int64 array gets 1-based entries, zeros past its count, and
// counts[a] its length; a row longer than maxneigh sets flag[0].  full: the
// rebuild-time full rows of full_rows_tile; half: the pairs of the stored
// short list with a < b, in list order.
template <int DIMS>
__device__ void ref_rows_tile(const Rep &R, const Tile &t, int half, long long *rows, long long *counts,
                              int maxneigh, int *flag)
{
    const int n = R.n, k = t.first + threadIdx.x;
    if (k >= n) return;
    const int a = gp(R.orig)[k];
    SC_AS1 long long *row = gp(rows) + (size_t)a * maxneigh;
    int cnt = 0;
    if (half) {
        const int m = gp(R.nn)[k];
        for (int q = 0; q < m; q++) {
            const int b = gp(R.orig)[gp(R.nbr)[(size_t)q * n + k]];
            if (a < b) {
                if (cnt < maxneigh) row[cnt] = (long long)b + 1;
                cnt++;
            }
        }
    } else {
        const double4 p = ld4(gp(R.pos_alt) + (k));
        int cc[3];
        cell_of_pos<DIMS>(R, p, cc);
        const int s = R.srange;
        const int zr = DIMS == 3 ? s : 0;
        for (int oz = -zr; oz <= zr; oz++)
            for (int oy = -s; oy <= s; oy++)
                for (int ox = -s; ox <= s; ox++) {
                    const int c = wrapc(cc[0] + ox, R.nc) +
                                  R.nc * (wrapc(cc[1] + oy, R.nc) + R.nc * (DIMS == 3 ? wrapc(cc[2] + oz, R.nc) : 0));
                    for (int j = cstart(R, c), je = cstart(R, c + 1); j < je; j++) {
                        if (j == k) continue;
                        double dx, dy, dz;
                        if (sqdist<DIMS>(R, p, ld4(gp(R.pos_alt) + (j)), dx, dy, dz) <= R.Rv2) {
                            if (cnt < maxneigh) row[cnt] = (long long)gp(R.orig)[j] + 1;
                            cnt++;
                        }
                    }
                }
    }
    for (int q = cnt; q < maxneigh; q++) row[q] = 0;
    gp(counts)[a] = cnt;
    if (cnt > maxneigh) gatomic_or(gp(flag), 1);
}

__global__ __launch_bounds__(SC_TB) void k_ref_rows(const Tile *tiles, const Rep *reps, int half, long long *rows,
                                                    long long *counts, int maxneigh, int *flag)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (R.dims == 3)
        ref_rows_tile<3>(R, t, half, rows, counts, maxneigh, flag);
    else
        ref_rows_tile<2>(R, t, half, rows, counts, maxneigh, flag);
}

// Fill kernel (no memset node, so a captured launch sequence replays it).
__global__ void k_zero_ints(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gp(p)[i] = 0;
}

// ====================================================================
// Energies
// ====================================================================

// calculate_full_energy (src/AuxiliaryFunctions.jl:22-34): per-particle sums
// over every candidate of the rebuild-time cell stencil with the inner loop's
// cutoff test, which selects exactly the full-list pairs inside the cutoff;
// Neumaier-compensated tile and replica reductions; total halved.  per: also
// compute_potential_energy! (src/AuxiliaryFunctions.jl:186-196), E_i / 2 of
// every particle at its original index.
template <int DIMS, int POT>
__device__ void energy_tile(const Rep &R, const Tile &t, double *per)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int n = R.n, k = t.first + threadIdx.x;
    double e = 0.0;
    if (k < n) {
        const double4 p = ld4(gp(R.pos) + (k));
        int cc[3];
        cell_of_pos<DIMS>(R, p, cc);
        const int s = R.srange;
        const int zr = DIMS == 3 ? s : 0;
        for (int oz = -zr; oz <= zr; oz++)
            for (int oy = -s; oy <= s; oy++)
                for (int ox = -s; ox <= s; ox++) {
                    const int c = wrapc(cc[0] + ox, R.nc) +
                                  R.nc * (wrapc(cc[1] + oy, R.nc) + R.nc * (DIMS == 3 ? wrapc(cc[2] + oz, R.nc) : 0));
                    for (int j = cstart(R, c), je = cstart(R, c + 1); j < je; j++) {
                        if (j == k) continue;
                        const double4 q = ld4(gp(R.pos) + (j));
                        double dx, dy, dz;
                        double d2 = sqdist<DIMS>(R, p, q, dx, dy, dz);
                        e += pair_energy<DIMS, POT>(R, d2, p.w, q.w);
                    }
                }
        if (per) gp(per)[gp(R.orig)[k]] = e / 2.0;
    }
    // compensated block reduction: exact two-sum tree is overkill here; a
    // per-wave Neumaier pass over lane values keeps the error at O(eps).
    __shared__ double ss[SC_TB];
    ss[threadIdx.x] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0, comp = 0.0;
        for (int q = 0; q < SC_TB; q++) neumaier_add(sum, comp, ss[q]);
        gp(R.part)[PS_ESUM * R.ntiles + t.local] = sum;
        gp(R.part)[PS_ECOMP * R.ntiles + t.local] = comp;
    }
}

__global__ __launch_bounds__(SC_TB) void k_energy(const Tile *tiles, const Rep *reps, double *per)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int key = (R.dims == 3 ? 2 : 0) + (R.pot == 2 ? 1 : 0);
    switch (key) {
    case 0: energy_tile<2, 1>(R, t, per); break;
    case 1: energy_tile<2, 2>(R, t, per); break;
    case 2: energy_tile<3, 1>(R, t, per); break;
    default: energy_tile<3, 2>(R, t, per); break;
    }
}

// calculate_full_energy_no_neigh (src/AuxiliaryFunctions.jl:86-110): all
// pairs, j tiles staged through LDS.
template <int DIMS, int POT>
__device__ void energy_nn_tile(const Rep &R, const Tile &t)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    __shared__ double4 sp[SC_TB];
    const int n = R.n, k = t.first + threadIdx.x;
    const double4 p = k < n ? ld4(gp(R.pos) + (k)) : make_double4(0, 0, 0, 1);
    double e = 0.0;
    for (int b = 0; b < n; b += SC_TB) {
        __syncthreads();
        if (b + threadIdx.x < n) sp[threadIdx.x] = ld4(gp(R.pos) + (b + threadIdx.x));
        __syncthreads();
        const int m = min(SC_TB, n - b);
        if (k < n)
            for (int q = 0; q < m; q++) {
                if (b + q == k) continue;
                double dx, dy, dz;
                double d2 = sqdist<DIMS>(R, p, sp[q], dx, dy, dz);
                e += pair_energy_generic<POT>(R, d2, p.w, sp[q].w);
            }
    }
    __shared__ double ss[SC_TB];
    ss[threadIdx.x] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0, comp = 0.0;
        for (int q = 0; q < SC_TB; q++) neumaier_add(sum, comp, ss[q]);
        gp(R.part)[PS_ESUM * R.ntiles + t.local] = sum;
        gp(R.part)[PS_ECOMP * R.ntiles + t.local] = comp;
    }
}

__global__ __launch_bounds__(SC_TB) void k_energy_nn(const Tile *tiles, const Rep *reps)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int key = (R.dims == 3 ? 2 : 0) + (R.pot == 2 ? 1 : 0);
    switch (key) {
    case 0: energy_nn_tile<2, 1>(R, t); break;
    case 1: energy_nn_tile<2, 2>(R, t); break;
    case 2: energy_nn_tile<3, 1>(R, t); break;
    default: energy_nn_tile<3, 2>(R, t); break;
    }
}

// find_kinetic_energy (src/AuxiliaryFunctions.jl:57-65); with apply != 0 it
// also stages rescale_velocities! (src/MD.jl:391-402) as a pending scale.
__global__ __launch_bounds__(SC_TB) void k_ekin(const Tile *tiles, const Rep *reps)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int n = R.n, i = t.first + threadIdx.x;
    double e[1] = {0.0};
    if (i < n) {
        double s = 0.0;
        for (int c = 0; c < R.dims; c++) {
            double vc = gp(R.v)[c * n + i];
            s = (c == 0) ? vc * vc : s + vc * vc;
        }
        e[0] = s;
    }
    block_sum<1>(e);
    put_partial(R, t, PS_SCAL, e[0]);
}

// ====================================================================
// Log-step observables: calculate_F2_self! and the Steinhardt order
// parameters (src/AuxiliaryFunctions.jl:687-902, 968-987)
// ====================================================================

// Compensated tile partial of one value per slot (PS_ESUM / PS_ECOMP).
__device__ inline void tile_partial_compensated(const Rep &R, const Tile &t, double e)
{
    __shared__ double ss[SC_TB];
    ss[threadIdx.x] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0, comp = 0.0;
        for (int q = 0; q < SC_TB; q++) neumaier_add(sum, comp, ss[q]);
        gp(R.part)[PS_ESUM * R.ntiles + t.local] = sum;
        gp(R.part)[PS_ECOMP * R.ntiles + t.local] = comp;
    }
}

// r_old_array .= r_array (src/MD.jl:682), kept in original particle order
// (component c of particle o at rold[c n + o]) so that it survives the
// re-sorting of slots at every rebuild.
__global__ __launch_bounds__(SC_TB) void k_set_rold(const Tile *tiles, const Rep *reps, double *rold)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int n = R.n, k = t.first + threadIdx.x;
    if (k >= n) return;
    const double4 p = ld4(gp(R.pos) + k);
    const int o = gp(R.orig)[k];
    SC_AS1 double *ro = gp(rold);
    ro[o] = p.x;
    ro[n + o] = p.y;
    if (R.dims == 3) ro[2 * n + o] = p.z;
}

// calculate_F2_self! (src/AuxiliaryFunctions.jl:968-987): per particle
// cos(k x_old) cos(k x) + sin(k x_old) sin(k x) on the x components.
__global__ __launch_bounds__(SC_TB) void k_f2s(const Tile *tiles, const Rep *reps, const double *rold, double kx)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int n = R.n, k = t.first + threadIdx.x;
    double e = 0.0;
    if (k < n) {
        const double x = ld4(gp(R.pos) + k).x;
        const double xo = gp(rold)[gp(R.orig)[k]];
        double so, co, s, c;
        sincos(kx * xo, &so, &co);
        sincos(kx * x, &s, &c);
        e = co * c + so * s;
    }
    tile_partial_compensated(R, t, e);
}

// Spherical harmonics Y_l^m(theta, phi), m = 0..L, orthonormal with the
// Condon-Shortley phase: K_l^m P_l^m(cos theta) e^{i m phi} with
// P_m^m = (-1)^m (2m-1)!! sin^m(theta) (the signed sine, so theta outside
// [0, pi] continues the formula), the upward recursion
// (l-m) P_l^m = (2l-1) x P_{l-1}^m - (l+m-1) P_{l-2}^m, and
// K_l^m = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!).  Added to (re, im).
template <int L>
__device__ inline void ylm_add(double theta, double phi, double (&re)[L + 1], double (&im)[L + 1])
{
    double s, x, sp, cp;
    sincos(theta, &s, &x);
    sincos(phi, &sp, &cp);
    double pmm = 1.0, cm = 1.0, sm = 0.0;
#pragma unroll
    for (int m = 0; m <= L; m++) {
        if (m > 0) {
            pmm *= -(2.0 * m - 1.0) * s;
            const double c2 = cm * cp - sm * sp;
            sm = sm * cp + cm * sp;
            cm = c2;
        }
        double plm = pmm;
        if (L > m) {
            double p0 = pmm, p1 = x * (2.0 * m + 1.0) * pmm;
#pragma unroll
            for (int ll = m + 2; ll <= L; ll++) {
                const double p2 = ((2.0 * ll - 1.0) * x * p1 - (ll + m - 1.0) * p0) / (double)(ll - m);
                p0 = p1;
                p1 = p2;
            }
            plm = p1;
        }
        double ratio = 1.0;
#pragma unroll
        for (int q = L - m + 1; q <= L + m; q++) ratio /= (double)q;
        const double K = sqrt((2.0 * L + 1.0) / (4.0 * M_PI) * ratio);
        re[m] += K * plm * cm;
        im[m] += K * plm * sm;
    }
}

// Visits every other slot j within the rebuild-time cell stencil of the
// particle at p whose minimum-image d2 <= qc2, as f(j, dx, dy, dz).
template <class Fn>
__device__ inline void qcut_neighbours(const Rep &R, int k, const double4 p, double qc2, Fn &&fn)
{
    int cc[3];
    cell_of_pos<3>(R, p, cc);
    const int s = R.srange, nc = R.nc;
    for (int oz = -s; oz <= s; oz++)
        for (int oy = -s; oy <= s; oy++)
            for (int ox = -s; ox <= s; ox++) {
                const int c = wrapc(cc[0] + ox, nc) + nc * (wrapc(cc[1] + oy, nc) + nc * wrapc(cc[2] + oz, nc));
                for (int j = cstart(R, c), je = cstart(R, c + 1); j < je; j++) {
                    if (j == k) continue;
                    const double4 q = ld4(gp(R.pos) + j);
                    const double dx = min_image(p.x - q.x, R.L, R.invL);
                    const double dy = min_image(p.y - q.y, R.L, R.invL);
                    const double dz = min_image(p.z - q.z, R.L, R.invL);
                    if (dx * dx + dy * dy + dz * dz > qc2) continue;
                    fn(j, dx, dy, dz);
                }
            }
}

// The reference's angles (src/AuxiliaryFunctions.jl:785-790): theta =
// atan(sqrt(x^2+y^2)/z), phi = atan(y/x) after nudging x == 0 by the
// literal 10^-10 (in Julia inv(10)^10).
constexpr double QL_NUDGE = 1.0000000000000006e-10;

// calculate_qlm for every m and calculate_ql's per-particle value
// (src/AuxiliaryFunctions.jl:687-711, 758-801): q_lm = mean over neighbours
// within q_cutoff of Y_l^m, stored (real at qlm[2m n + k], imaginary at
// qlm[(2m+1) n + k], count at nq[k]) for the averaged form; the tile
// partials sum q_l(i) = sqrt(4 pi/(2l+1) sum_{m=-l..l} |q_lm|^2), using
// |q_{l,-m}| = |q_lm|.
template <int L>
__global__ __launch_bounds__(SC_TB) void k_qlm(const Tile *tiles, const Rep *reps, double qc2, double *qlm, int *nq)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int n = R.n, k = t.first + threadIdx.x;
    double e = 0.0;
    if (k < n) {
        double re[L + 1], im[L + 1];
#pragma unroll
        for (int m = 0; m <= L; m++) re[m] = im[m] = 0.0;
        int cnt = 0;
        const double4 p = ld4(gp(R.pos) + k);
        qcut_neighbours(R, k, p, qc2, [&](int, double dx, double dy, double dz) {
            cnt++;
            const double theta = atan(sqrt(dx * dx + dy * dy) / dz);
            const double phi = atan(dy / (dx == 0.0 ? dx + QL_NUDGE : dx));
            ylm_add<L>(theta, phi, re, im);
        });
        double tot = 0.0;
#pragma unroll
        for (int m = 0; m <= L; m++) {
            if (cnt > 0) {
                re[m] /= (double)cnt;
                im[m] /= (double)cnt;
            }
            gp(qlm)[(size_t)(2 * m) * n + k] = re[m];
            gp(qlm)[(size_t)(2 * m + 1) * n + k] = im[m];
            const double a = hypot(re[m], im[m]);
            tot += (m == 0 ? 1.0 : 2.0) * (a * a);
        }
        gp(nq)[k] = cnt;
        e = sqrt(4.0 * M_PI / (2.0 * L + 1.0) * tot);
    }
    tile_partial_compensated(R, t, e);
}

// calculate_ql_averaged's second pass (src/AuxiliaryFunctions.jl:866-893):
// (q_lm(i) + sum over the same neighbours of q_lm(j)) / (N_q(i) + 1).
template <int L>
__global__ __launch_bounds__(SC_TB) void k_qlm_avg(const Tile *tiles, const Rep *reps, double qc2,
                                                  const double *qlm, const int *nq)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int n = R.n, k = t.first + threadIdx.x;
    double e = 0.0;
    if (k < n) {
        double re[L + 1], im[L + 1];
        const SC_AS1 double *Q = gp(qlm);
#pragma unroll
        for (int m = 0; m <= L; m++) {
            re[m] = Q[(size_t)(2 * m) * n + k];
            im[m] = Q[(size_t)(2 * m + 1) * n + k];
        }
        const double4 p = ld4(gp(R.pos) + k);
        qcut_neighbours(R, k, p, qc2, [&](int j, double, double, double) {
#pragma unroll
            for (int m = 0; m <= L; m++) {
                re[m] += Q[(size_t)(2 * m) * n + j];
                im[m] += Q[(size_t)(2 * m + 1) * n + j];
            }
        });
        const double den = (double)(gp(nq)[k] + 1);
        double tot = 0.0;
#pragma unroll
        for (int m = 0; m <= L; m++) {
            const double a = hypot(re[m] / den, im[m] / den);
            tot += (m == 0 ? 1.0 : 2.0) * (a * a);
        }
        e = sqrt(tot * (4.0 * M_PI) / (2.0 * L + 1.0));
    }
    tile_partial_compensated(R, t, e);
}

// ====================================================================
// Host-order transfers: staging buffer indexed by original particle index.
// ====================================================================
__global__ __launch_bounds__(SC_TB) void k_to_host(const Tile *tiles, const Rep *reps, int field)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int n = R.n, k = t.first + threadIdx.x, d = R.dims;
    if (k >= n) return;
    const int o = gp(R.orig)[k];
    double *st = R.stage;
    switch (field) {
    case FLD_R: {
        double4 p = ld4(gp(R.pos) + (k));
        st[(size_t)o * d + 0] = p.x;
        st[(size_t)o * d + 1] = p.y;
        if (d == 3) st[(size_t)o * d + 2] = p.z;
    } break;
    case FLD_D: st[o] = ld4(gp(R.pos) + (k)).w; break;
    case FLD_V:
    case FLD_F:
    case FLD_DISP: {
        const double *src = field == FLD_V ? R.v : (field == FLD_F ? R.F : R.disp);
        for (int c = 0; c < d; c++) st[(size_t)o * d + c] = src[c * n + k];
    } break;
    case FLD_DISP2: {
        double s = 0.0;
        for (int c = 0; c < d; c++) {
            double x = gp(R.disp)[c * n + k];
            s = (c == 0) ? x * x : s + x * x;
        }
        st[o] = s;
    } break;
    }
}

__global__ __launch_bounds__(SC_TB) void k_from_host(const Tile *tiles, const Rep *reps, int field)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    const int n = R.n, k = t.first + threadIdx.x, d = R.dims;
    if (k >= n) return;
    const int o = gp(R.orig)[k];
    const double *st = R.stage;
    switch (field) {
    case FLD_R: {
        // stored as the periodic image in [0, L), with the drift's wrap
        // (src/MD.jl:520): the cell binning and the builds' image shifts
        // take wrapped positions, and the reference's lists and forces
        // depend on positions only modulo the box
        double4 p = ld4(gp(R.pos) + (k));
        double x[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < d; c++) {
            x[c] = st[(size_t)o * d + c];
            x[c] -= floor(x[c] / R.L) * R.L;
        }
        p.x = x[0];
        p.y = x[1];
        p.z = x[2];
        st4(gp(R.pos) + (k), p);
        put_shadow(R, k, p);
    } break;
    case FLD_D: {
        double4 p = ld4(gp(R.pos) + (k));
        p.w = st[o];
        st4(gp(R.pos) + (k), p);
        put_shadow(R, k, p);
    } break;
    case FLD_V:
    case FLD_F:
    case FLD_DISP: {
        double *dst = field == FLD_V ? R.v : (field == FLD_F ? R.F : R.disp);
        for (int c = 0; c < d; c++) dst[c * n + k] = st[(size_t)o * d + c];
    } break;
    }
}

// ====================================================================
// Capped steepest descent (find_inherent_structure,
// src/InherentStructure.jl:26-51)
// ====================================================================
__global__ __launch_bounds__(SC_TB) void k_min_scale(const Tile *tiles, const Rep *reps, double dt)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    if (R.min_tol >= 0.0 && dyn->min_state != 0) return;  // converged (tolerance run)
    const int n = R.n, i = t.first + threadIdx.x;
    double st = 0.0, fm = 0.0;
    if (i < n) {
        double s = 0.0, f2 = 0.0;
        for (int c = 0; c < R.dims; c++) {
            double fc = gp(R.F)[c * n + i];
            double vc = fc * dt;
            s = (c == 0) ? vc * vc : s + vc * vc;
            f2 = (c == 0) ? fc * fc : f2 + fc * fc;
        }
        st = sqrt(s);
        fm = sqrt(f2);
    }
    double m1 = block_max(st);
    double m2 = block_max(fm);
    put_partial(R, t, PS_AUXMAX, m1);
    put_partial(R, t, PS_VEC, m2);
}

template <int DIMS>
__device__ void min_move_tile(const Rep &R, const Tile &t, double dt)
{
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    if (R.min_tol >= 0.0 && dyn->min_state == 2) return;  // converged and moved (tolerance run)
    const int n = R.n, i = t.first + threadIdx.x;
    const double scale = dyn->result[2];
    double s2 = 0.0;
    if (i < n) {
        double4 p = ld4(gp(R.pos) + (i));
        double pc[3] = {p.x, p.y, p.z};
        double s = 0.0;
        bool fin = true;
#pragma unroll
        for (int c = 0; c < DIMS; c++) {
            double vc = gp(R.F)[c * n + i] * dt;
            if (scale >= 0.0) vc = vc * scale;
            gp(R.v)[c * n + i] = vc;
            double dc = gp(R.disp)[c * n + i] + vc;
            s = (c == 0) ? dc * dc : s + dc * dc;
            double rn = pc[c] + vc;
            rn -= floor(rn / R.L) * R.L;
            pc[c] = rn;
            fin = fin && isfinite(rn);
            gp(R.disp)[c * n + i] = dc;
        }
        const double4 pn = make_double4(pc[0], pc[1], pc[2], p.w);
        st4(gp(R.pos) + (i), pn);
        put_shadow(R, i, pn);
        s2 = s;
        if (!fin) gatomic_or(&dyn->err, SC_EBIT_NONFINITE);
    }
    // F stays live across the rebuild: the forces of the last iteration are
    // the minimiser's output beside the positions (src/InherentStructure.jl:25)
    publish_maxdisp(R, block_max(s2), 1 | REBUILD_KEEP_F);
}

__global__ __launch_bounds__(SC_TB) void k_min_move(const Tile *tiles, const Rep *reps, double dt)
{
    const Tile t = tiles[xcd_tile()];
    const Rep R = reps[t.rep];
    if (R.dims == 3)
        min_move_tile<3>(R, t, dt);
    else
        min_move_tile<2>(R, t, dt);
}

// ====================================================================
// k_finish: one block per replica turns the tile partials of the previous
// launch into per-replica results (fixed order => deterministic).
// ====================================================================
enum {
    FIN_DRIFT = 0,      // max displacement^2 -> rebuild flag; pending COM consumed
    FIN_FORCE0,         // force evaluation counter
    FIN_FORCE1,         // + COM velocity (pending)
    FIN_FORCE2,         // + mean force (Brownian)
    FIN_EKCOM,          // rescale factor from the COM-shifted kinetic energy
    FIN_FINALIZE,       // pending COM applied
    FIN_REBUILD_DONE,   // rebuild bookkeeping
    FIN_ENERGY,         // compensated energy total, halved
    FIN_EKIN,           // kinetic energy
    FIN_EKIN_RESCALE,   // + rescale_velocities! factor (pending)
    FIN_MINSCALE,       // minimiser step scale and max |F|
    FIN_MIN_BEGIN,      // minimiser tolerance run: reset its state and count
    FIN_MIN_END,        // ... and leave the force kernels unblocked
    FIN_MEAN,           // compensated total / n
    FIN_FORCE1_AHEAD,   // FIN_FORCE1 of a step launched ahead: halts the chunk instead when the trigger fired
    FIN_FORCE2_AHEAD,   // the same for FIN_FORCE2
    FIN_FORCE1_RESUME,  // FIN_FORCE1 of the resumed step of a halted replica (others: nothing)
    FIN_SET_TARGET      // target = step + arg (arg >= 0), else none
};

// Sums of K tile-partial slots over a replica's nt tiles: thread t adds the
// tiles t, t + SC_TB, t + 2 SC_TB, ... in that order (the order fixes the
// bits), with the loads of up to eight such tiles issued before their adds
// (the partials were written by blocks on every XCD and come from memory).
template <int K>
__device__ inline void tile_sums(const SC_AS1 double *P, const int (&slot)[K], int nt, double (&a)[K])
{
    constexpr int U = 8;
    for (int q0 = threadIdx.x; q0 < nt; q0 += U * SC_TB) {
        double v[U][K];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = q0 + u * SC_TB;
#pragma unroll
            for (int c = 0; c < K; c++) v[u][c] = q < nt ? P[(size_t)slot[c] * nt + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (q0 + u * SC_TB >= nt) break;
#pragma unroll
            for (int c = 0; c < K; c++) a[c] += v[u][c];
        }
    }
}

__global__ __launch_bounds__(SC_TB) void k_finish(const Rep *reps, int nreps, int op, int sysmask, double arg)
{
    const int r = blockIdx.x;
    if (r >= nreps) return;
    const Rep R = reps[r];
    if (!((sysmask >> R.sys) & 1)) return;
    SC_AS1 RepDyn *dyn = gp(R.dyn);
    const int tid = threadIdx.x;
    const SC_AS1 double *P = gp(R.part);
    if (op == FIN_FORCE1_AHEAD || op == FIN_FORCE2_AHEAD) {
        const int h = dyn->halt, rb = dyn->rebuild;
        const bool done = dyn->step >= dyn->target;
        __syncthreads();  // every thread has read the flags before thread 0 may set halt
        if (done) {
            // the replica's steps of this call are done (its force kernel
            // returned too): bit 1 stops the rest of this step (the Brownian
            // drift) and every later step of the call
            if (tid == 0) dyn->halt = h | 2;
            return;
        }
        if (h || rb) {
            if (!h && tid == 0) dyn->halt = 1;  // the host runs the rebuild and this step again
            return;
        }
        op = op == FIN_FORCE1_AHEAD ? FIN_FORCE1 : FIN_FORCE2;
    } else if (op == FIN_FORCE1_RESUME) {
        const int h = dyn->halt;
        __syncthreads();  // read before FIN_FORCE1 clears it
        if (!(h & 1)) return;
        op = FIN_FORCE1;
    } else if (op == FIN_SET_TARGET) {
        if (tid == 0) {
            dyn->target = arg >= 0.0 ? dyn->step + (long long)arg : LLONG_MAX;
            dyn->halt &= 1;  // a finished call's done bit
        }
        return;
    }
    switch (op) {
    case FIN_DRIFT:
        break;
    case FIN_FORCE0:
        if (tid == 0) {
            // a stopped minimiser's force kernel returned at once: no
            // evaluation to count (the step counter keys the noise streams)
            if (R.min_tol < 0.0 || dyn->min_state == 0) dyn->step += 1;
            else if (dyn->min_state == 1) dyn->min_state = 2;  // the converged iteration's move is done
        }
        break;
    case FIN_MIN_BEGIN:
        if (tid == 0) {
            dyn->min_state = 0;
            dyn->min_iters = 0;
        }
        break;
    case FIN_MIN_END:
        if (tid == 0) dyn->min_state = 0;
        break;
    case FIN_FORCE1:
    case FIN_FORCE2: {
        double a[3] = {0.0, 0.0, 0.0};
        const int slots[3] = {PS_VEC, PS_VEC + 1, PS_VEC + 2};
        tile_sums<3>(P, slots, R.ntiles, a);
        block_sum<3>(a);
        if (tid == 0) {
            for (int c = 0; c < R.dims; c++) {
                double mean = a[c] / (double)R.n;
                if (op == FIN_FORCE1)
                    dyn->vmean[c] = mean;
                else
                    dyn->Fmean[c] = mean;
            }
            if (op == FIN_FORCE1) {
                // rebuild bookkeeping of this step (the rebuild kernels ran
                // between the drift and this force evaluation)
                if (dyn->rebuild) {
                    dyn->rebuilds += 1;
                    dyn->have_lists = 1;
                }
                dyn->rebuild = 0;
                dyn->halt = 0;
                dyn->pending = 1 | (R.rescale ? 2 : 0);
            } else if (dyn->rebuild) {
                // Brownian: bookkeeping of the previous step's rebuild (which
                // ran after its drift; this force already used the new lists),
                // before this step's drift raises the flag again
                dyn->rebuilds += 1;
                dyn->have_lists = 1;
                dyn->rebuild = 0;
            }
            dyn->step += 1;
        }
    } break;
    case FIN_EKCOM:
    case FIN_EKIN:
    case FIN_EKIN_RESCALE: {
        double a[1] = {0.0};
        const int slots[1] = {PS_SCAL};
        tile_sums<1>(P, slots, R.ntiles, a);
        block_sum<1>(a);
        if (tid == 0) {
            double ek = a[0] * (R.m / 2.0);
            double kT = 2.0 * ek / ((double)R.dims * (double)R.n);
            if (op == FIN_EKCOM) {
                dyn->vscale = sqrt(R.kBT / kT);
            } else {
                dyn->result[1] = ek;
                if (op == FIN_EKIN_RESCALE) {
                    dyn->vscale = sqrt(R.kBT / kT);
                    dyn->pending = 2;
                }
            }
        }
    } break;
    case FIN_FINALIZE:
        if (tid == 0) dyn->pending = 0;
        break;
    case FIN_REBUILD_DONE:
        if (tid == 0) {
            if (dyn->rebuild) {
                dyn->rebuilds += 1;
                dyn->have_lists = 1;
            }
            dyn->rebuild = 0;
            dyn->halt = 0;
            dyn->pending = 0;
        }
        break;
    case FIN_ENERGY:
        if (tid == 0) {
            double sum = 0.0, comp = 0.0;
            for (int q = 0; q < R.ntiles; q++) {
                neumaier_add(sum, comp, P[PS_ESUM * R.ntiles + q]);
                comp += P[PS_ECOMP * R.ntiles + q];
            }
            dyn->result[0] = (sum + comp) / 2.0;
        }
        break;
    case FIN_MEAN:
        if (tid == 0) {
            double sum = 0.0, comp = 0.0;
            for (int q = 0; q < R.ntiles; q++) {
                neumaier_add(sum, comp, P[PS_ESUM * R.ntiles + q]);
                comp += P[PS_ECOMP * R.ntiles + q];
            }
            dyn->result[0] = (sum + comp) / (double)R.n;
        }
        break;
    case FIN_MINSCALE: {
        if (R.min_tol >= 0.0 && dyn->min_state != 0) break;  // converged: the iteration is a no-op
        double st = 0.0, fm = 0.0;
        for (int q = tid; q < R.ntiles; q += SC_TB) {
            st = fmax(st, P[PS_AUXMAX * R.ntiles + q]);
            fm = fmax(fm, P[PS_VEC * R.ntiles + q]);
        }
        st = block_max(st);
        fm = block_max(fm);
        if (tid == 0) {
            dyn->result[2] = (st > arg) ? arg / st : -1.0;  // -1: unscaled
            dyn->result[3] = fm;
            dyn->min_iters += 1;
            // src/InherentStructure.jl:50-54: error = max(max|F|, step) with the
            // step taken before the cap (= max|F| dt < max|F|), checked after the move
            if (R.min_tol >= 0.0 && fmax(fm, st) <= R.min_tol) dyn->min_state = 1;
        }
    } break;
    }
}

inline int check(hipError_t e) { return e == hipSuccess ? 0 : -5; }

}  // namespace

// ====================================================================
// Launchers: each tile kernel is followed by its k_finish.
// ====================================================================
namespace sc {

constexpr int SYS_ALL = 7, SYS_VV = 3, SYS_BD = 4, SYS_NEWT = 1;

static inline void fin(const Rep *reps, int nreps, int op, int mask, double arg, hipStream_t s)
{
    hipLaunchKernelGGL(k_finish, dim3(nreps), dim3(SC_TB), 0, s, reps, nreps, op, mask, arg);
}

// Plain streaming copy (the measured HBM ceiling of the measurement hook).
__global__ __launch_bounds__(SC_TB) void k_copy(const double4 *src, double4 *dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)SC_TB + threadIdx.x; i < n; i += (size_t)gridDim.x * SC_TB)
        st4(gp(dst) + i, ld4(gp(src) + i));
}

int launch_copy(const double4 *src, double4 *dst, size_t n, hipStream_t s)
{
    const size_t blocks = (n + SC_TB - 1) / SC_TB;
    hipLaunchKernelGGL(k_copy, dim3(blocks < 8192 ? (unsigned)blocks : 8192u), dim3(SC_TB), 0, s, src, dst, n);
    return check(hipGetLastError());
}

int launch_kick_drift(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_kick_drift, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, 0);
    return check(hipGetLastError());
}

int launch_force(const Tile *tiles, int ntiles, const Rep *reps, int nreps, int mode, int sysmask, hipStream_t s,
                 bool finish, bool ahead)
{
    if (ntiles <= 0) return 0;
    const int a = ahead ? 1 : 0;
    if (mode == 0)
        hipLaunchKernelGGL(k_force<0>, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, sysmask, a);
    else if (mode == 1)
        hipLaunchKernelGGL(k_force<1>, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, sysmask, a);
    else
        hipLaunchKernelGGL(k_force<2>, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, sysmask, a);
    if (finish) return launch_force_finish(reps, nreps, mode, sysmask, s, ahead);
    return check(hipGetLastError());
}

int launch_force_finish(const Rep *reps, int nreps, int mode, int sysmask, hipStream_t s, bool ahead)
{
    const int op = mode == 0 ? FIN_FORCE0
                             : (mode == 1 ? (ahead ? FIN_FORCE1_AHEAD : FIN_FORCE1) : (ahead ? FIN_FORCE2_AHEAD : FIN_FORCE2));
    fin(reps, nreps, op, sysmask, 0.0, s);
    return check(hipGetLastError());
}

int launch_set_target(const Rep *reps, int nreps, long long nsteps, hipStream_t s)
{
    fin(reps, nreps, FIN_SET_TARGET, SYS_ALL, nsteps >= 0 ? (double)nsteps : -1.0, s);
    return check(hipGetLastError());
}

int launch_force_resume(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_force<1>, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, 3, 2);
    fin(reps, nreps, FIN_FORCE1_RESUME, 3, 0.0, s);
    return check(hipGetLastError());
}

int launch_kick_drift_only(const Tile *tiles, int ntiles, const Rep *reps, hipStream_t s, bool ahead)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_kick_drift, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, ahead ? 1 : 0);
    return check(hipGetLastError());
}

int launch_drift_brownian(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s, bool ahead)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_drift_brownian, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, ahead ? 1 : 0);
    return check(hipGetLastError());
}

int launch_ekin_com(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_ekin_com, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps);
    fin(reps, nreps, FIN_EKCOM, SYS_NEWT, 0.0, s);
    return check(hipGetLastError());
}

int launch_rebuild_done(const Rep *reps, int nreps, hipStream_t s, int sysmask)
{
    fin(reps, nreps, FIN_REBUILD_DONE, sysmask, 0.0, s);
    return check(hipGetLastError());
}

int launch_request_rebuild(const Rep *reps, int nreps, hipStream_t s)
{
    hipLaunchKernelGGL(k_request_rebuild, dim3((nreps + 255) / 256), dim3(256), 0, s, const_cast<Rep *>(reps), nreps);
    return check(hipGetLastError());
}

int launch_gather_err(const Rep *reps, int nreps, int *out, hipStream_t s, int what)
{
    if (nreps <= 0) return 0;
    hipLaunchKernelGGL(k_gather_err, dim3((nreps + SC_TB - 1) / SC_TB), dim3(SC_TB), 0, s, reps, nreps, out, what);
    return check(hipGetLastError());
}

bool build_shape_valid(int bs, int lpp)
{
    if (lpp == 0) return bs == 256;  // the per-slot radius form
    return (bs == 64 || bs == 256) && (lpp == 1 || lpp == 2 || lpp == 4 || lpp == 8 || lpp == 16 || lpp == 32);
}

template <int BS, int LPP>
static void launch_build_shape(const Tile *tiles, int ntiles, const Rep *reps, int nreps, const int *plan,
                               const int *flag, hipStream_t s, int which)
{
    const int items = ntiles * (SC_TB * LPP / BS);
    const int cap = SC_GRID_CAP * (SC_TB / BS);
    const dim3 grid(items < cap ? items : cap);
    if (which & 1) hipLaunchKernelGGL((k_build<BS, LPP>), grid, dim3(BS), 0, s, tiles, reps, plan, flag, nreps, ntiles);
    if (which & 2)
        hipLaunchKernelGGL((k_build_fx<BS, LPP>), grid, dim3(BS), 0, s, tiles, reps, plan, flag, nreps, ntiles);
}

int launch_rebuild(const Tile *tiles, int ntiles, const Rep *reps, int nreps, const RebuildGeom &g, hipStream_t s,
                   bool bookkeep)
{
    if (ntiles <= 0) return 0;
    if (!build_shape_valid(g.build_bs, g.build_lpp)) return -1;
    // one replica: no work list, every rebuild kernel checks the flag itself
    const int *plan = (nreps == 1 && g.flag) ? nullptr : g.plan;
    if (!plan && !g.flag) return -1;
    const int gt = ntiles < SC_GRID_CAP ? ntiles : SC_GRID_CAP;
    if (plan) hipLaunchKernelGGL(k_rebuild_plan, dim3(1), dim3(SC_TB), 0, s, reps, nreps, ntiles, g.plan);
    const int *flag = g.flag;
    const int fused = g.max_chunks <= SC_FUSED_CHUNKS ? 1 : 0;
    hipLaunchKernelGGL(k_bin, dim3(gt), dim3(SC_TB), 0, s, tiles, reps, plan, flag, nreps, ntiles, fused);
    hipLaunchKernelGGL(k_scan_local, dim3(g.max_chunks, nreps), dim3(SC_TB), 0, s, reps, plan, flag, fused);
    if (!fused) {
        hipLaunchKernelGGL(k_scan_top, dim3(nreps), dim3(SC_TB), 0, s, reps, plan, flag);
        hipLaunchKernelGGL(k_scan_fold, dim3(g.max_chunks, nreps), dim3(SC_TB), 0, s, reps, plan, flag);
    }
    hipLaunchKernelGGL(k_scatter, dim3(gt), dim3(SC_TB), 0, s, tiles, reps, plan, flag, nreps, ntiles, fused);
    hipLaunchKernelGGL(k_rank_permute, dim3(gt), dim3(SC_TB), 0, s, tiles, reps, plan, flag, nreps, ntiles);
    if (g.nruns > 0) {
        const int gr = g.nruns < SC_RUN_GRID_CAP ? g.nruns : SC_RUN_GRID_CAP;
        hipLaunchKernelGGL(k_build_cells, dim3(gr), dim3(SC_TB), 0, s, reps, nreps, g.nruns);
    }
    if (g.nwin > 0) {
        const int gw = g.nwin < SC_RUN_GRID_CAP ? g.nwin : SC_RUN_GRID_CAP;
        hipLaunchKernelGGL(k_build_win, dim3(gw), dim3(SC_TB), 0, s, reps, plan, flag, nreps, ntiles);
    }
    if (g.any_rows_build) {
        const int items = ntiles * (SC_TB / 64), cap = 4 * SC_GRID_CAP;
        const dim3 grid(items < cap ? items : cap);
        if (g.any_rows_build & 8)
            hipLaunchKernelGGL(k_build_rows<3>, grid, dim3(64), 0, s, tiles, reps, plan, flag, nreps, ntiles);
        if (g.any_rows_build & 4)
            hipLaunchKernelGGL(k_build_rows<2>, grid, dim3(64), 0, s, tiles, reps, plan, flag, nreps, ntiles);
    }
    if (g.any_tile_build) {
        const int key = g.build_bs * 64 + g.build_lpp, w = g.any_tile_build;
        switch (key) {
        case 256 * 64 + 0: {
            // per-slot radius form for the prefiltered replicas, the FP64 sweep for the others
            const dim3 grid(ntiles < SC_GRID_CAP ? ntiles : SC_GRID_CAP);
            if (w & 1) launch_build_shape<256, 1>(tiles, ntiles, reps, nreps, plan, flag, s, 1);
            if (w & 2) hipLaunchKernelGGL(k_build_px, grid, dim3(SC_TB), 0, s, tiles, reps, plan, flag, nreps, ntiles);
            break;
        }
        case 64 * 64 + 1: launch_build_shape<64, 1>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 64 * 64 + 2: launch_build_shape<64, 2>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 64 * 64 + 4: launch_build_shape<64, 4>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 64 * 64 + 8: launch_build_shape<64, 8>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 64 * 64 + 16: launch_build_shape<64, 16>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 64 * 64 + 32: launch_build_shape<64, 32>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 256 * 64 + 1: launch_build_shape<256, 1>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 256 * 64 + 2: launch_build_shape<256, 2>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 256 * 64 + 4: launch_build_shape<256, 4>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 256 * 64 + 8: launch_build_shape<256, 8>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        case 256 * 64 + 16: launch_build_shape<256, 16>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        default: launch_build_shape<256, 32>(tiles, ntiles, reps, nreps, plan, flag, s, w); break;
        }
    }
    if (bookkeep) fin(reps, nreps, FIN_REBUILD_DONE, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_finalize(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_finalize, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps);
    fin(reps, nreps, FIN_FINALIZE, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_energy(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s, double *per)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_energy, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, per);
    fin(reps, nreps, FIN_ENERGY, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_energy_nn(const Tile *tiles, int ntiles, const Rep *reps, int nreps, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_energy_nn, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps);
    fin(reps, nreps, FIN_ENERGY, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_set_rold(const Tile *tiles, int ntiles, const Rep *reps, double *rold, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_set_rold, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, rold);
    return check(hipGetLastError());
}

int launch_f2s(const Tile *tiles, int ntiles, const Rep *reps, int nreps, const double *rold, double kx,
               hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_f2s, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, rold, kx);
    fin(reps, nreps, FIN_MEAN, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

template <int L>
static void launch_ql_l(const Tile *tiles, int ntiles, const Rep *reps, double qc2, int averaged, double *qlm,
                        int *nq, hipStream_t s)
{
    hipLaunchKernelGGL(k_qlm<L>, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, qc2, qlm, nq);
    if (averaged) hipLaunchKernelGGL(k_qlm_avg<L>, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, qc2, qlm, nq);
}

int launch_steinhardt(const Tile *tiles, int ntiles, const Rep *reps, int nreps, int l, double qc2, int averaged,
                      double *qlm, int *nq, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    switch (l) {
    case 1: launch_ql_l<1>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 2: launch_ql_l<2>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 3: launch_ql_l<3>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 4: launch_ql_l<4>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 5: launch_ql_l<5>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 6: launch_ql_l<6>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 7: launch_ql_l<7>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    case 8: launch_ql_l<8>(tiles, ntiles, reps, qc2, averaged, qlm, nq, s); break;
    default: return -1;
    }
    fin(reps, nreps, FIN_MEAN, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_ekin(const Tile *tiles, int ntiles, const Rep *reps, int nreps, int apply_rescale, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_ekin, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps);
    fin(reps, nreps, apply_rescale ? FIN_EKIN_RESCALE : FIN_EKIN, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_from_host_order(const Tile *tiles, int ntiles, const Rep *reps, int field, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_from_host, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, field);
    return check(hipGetLastError());
}

int launch_to_host_order(const Tile *tiles, int ntiles, const Rep *reps, int field, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_to_host, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, field);
    return check(hipGetLastError());
}

int launch_ref_rows(const Tile *tiles, int ntiles, const Rep *reps, int half, long long *rows, long long *counts,
                    int maxneigh, int *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, flag, 1);
    hipLaunchKernelGGL(k_ref_rows, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, half, rows, counts, maxneigh, flag);
    return check(hipGetLastError());
}

int launch_full_rows(const Tile *tiles, int ntiles, const Rep *reps, int *rows, int *counts, int cap, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_full_rows, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, rows, counts, cap);
    return check(hipGetLastError());
}

int launch_min_scale(const Tile *tiles, int ntiles, const Rep *reps, int nreps, double dt, double max_step,
                     hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_min_scale, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, dt);
    fin(reps, nreps, FIN_MINSCALE, SYS_ALL, max_step, s);
    return check(hipGetLastError());
}

int launch_min_mark(const Rep *reps, int nreps, bool begin, hipStream_t s)
{
    fin(reps, nreps, begin ? FIN_MIN_BEGIN : FIN_MIN_END, SYS_ALL, 0.0, s);
    return check(hipGetLastError());
}

int launch_min_move(const Tile *tiles, int ntiles, const Rep *reps, int nreps, double dt, hipStream_t s)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_min_move, dim3(ntiles), dim3(SC_TB), 0, s, tiles, reps, dt);
    return check(hipGetLastError());
}

void philox_host(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) { philox4x32_10(in, key, out); }

}  // namespace sc
