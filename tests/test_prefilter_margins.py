"""The neighbour-list build's fixed-point FP32 prefilter, restated in numpy
float32 (IEEE round-to-nearest; the device code is compiled with
-ffp-contract=off, so every product and sum is rounded separately except the
explicit fused multiply-adds, restated here by fma32), checked against the
FP64 predicates of the reference on adversarial pairs.

The device test (simulationcode_amd/csrc/sc_kernels.hip, prefilter_q and
prefilter_half_bounds) is:
  u = rint(x * 2^32 / L) mod 2^32 per coordinate (FP64 product, then rint);
  d2 = sum of squares of float32(int32(u_i - u_j)), in FP32, left to right;
  full list: d2 > hi -> out, d2 < lo -> in, else the FP64 test;
  half list (Berthier): t = fma(-zeta, |Di - Dj|, 1), rh = fma(c1, (Di + Dj) t,
  c2) in FP32 from FP32 diameters, with c1 = rcut S / 2 and c2 = skin S;
  in if d2 < fma(rh^2, hlo, -m), out if d2 > fma(rh^2, hhi, m), else the
  FP64 test.
The margins (lo, hi, hlo, hhi, m) follow build_prefilter_setup
(simulationcode_amd/csrc/sc_replica.hip).  The test places pairs at the
FP64 thresholds and at relative offsets of 1e-9 .. 1e-3 around them, in
every direction and across the periodic boundary, and requires every
decision the prefilter takes without the FP64 test to equal the FP64
predicate (d2 <= R_v^2, src/NeighborLists.jl cutoff test; d2 < (rcut
sigma_ij + skin)^2, src/NeighborLists.jl:92).  A planted mistake (margins a
hundred times smaller) must fail, so the check has teeth.
"""
import numpy as np
import pytest


def fma32(a, b, c):
    """Correctly rounded float32 a*b + c.  The float64 product of two
    float32 values is exact; the float64 sum s carries the rounding error e
    (TwoSum, p + c == s + e exactly), and rounding s to float32 differs from
    rounding p + c only when s is a float32 midpoint and e != 0, where the
    result is the neighbour on e's side."""
    a, b, c = (np.asarray(x, dtype=np.float32).astype(np.float64) for x in (a, b, c))
    p = a * b
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    lo_n = np.nextafter(r, np.float32(-np.inf)).astype(np.float64)
    hi_n = np.nextafter(r, np.float32(np.inf)).astype(np.float64)
    mid = ((s - r64) * 2.0 == (hi_n - r64)) | ((r64 - s) * 2.0 == (r64 - lo_n))
    above = np.where(r64 > s, r, np.nextafter(r, np.float32(np.inf)))
    below = np.where(r64 < s, r, np.nextafter(r, np.float32(-np.inf)))
    r = np.where(mid & (e > 0.0), above, np.where(mid & (e < 0.0), below, r))
    return r.astype(np.float32)


def margins(L, Rv, dims, zeta, dmax, rcut, skin, lj_md=None, shrink=1.0):
    S = 2.0 ** 32 / L
    u = 2.0 ** -24
    RvS2 = Rv * Rv * S * S
    m4 = shrink * 4.0 * (6.0 * u * RvS2 + 2.1 * np.sqrt(dims) * Rv * S + 1.1 * dims)
    eh = shrink * 8.0 * (23.0 + 16.0 * abs(zeta) * dmax) * u
    out = dict(S=np.float32(S), lo=np.float32(RvS2 - m4), hi=np.float32(RvS2 + m4), m=np.float32(m4),
               hlo=np.float32(1.0 - eh), hhi=np.float32(1.0 + eh), c1=np.float32(0.5 * rcut * S),
               c2=np.float32(skin * S), zeta=np.float32(zeta))
    if lj_md is not None:
        rh = rcut * lj_md + skin
        rh2 = rh * rh * S * S
        out["rh2lo"] = np.float32(rh2 * (1.0 - 2.0 ** -20) - m4)
        out["rh2hi"] = np.float32(rh2 * (1.0 + 2.0 ** -20) + m4)
    return out


def fixed(x, L):
    return (np.rint(x * (2.0 ** 32 / L)).astype(np.int64) % (2 ** 32)).astype(np.uint32)


def prefilter(ri, rj, Di, Dj, L, M, berthier=True):
    """Device decisions: arrays (full, half, exact) where exact marks the pairs
    the prefilter hands to the FP64 test (their full/half entries are unset)."""
    ui, uj = fixed(ri, L), fixed(rj, L)
    du = (ui - uj).view(np.int32).astype(np.float32)
    d2 = du[:, 0] * du[:, 0] + du[:, 1] * du[:, 1]
    for c in range(2, ri.shape[1]):
        d2 = d2 + du[:, c] * du[:, c]
    full = d2 < M["lo"]
    exact = (d2 >= M["lo"]) & (d2 <= M["hi"])
    if berthier:
        si, sj = Di.astype(np.float32), Dj.astype(np.float32)
        t = fma32(-M["zeta"], np.abs(si - sj), np.float32(1.0))
        rh = fma32(M["c1"], (si + sj) * t, M["c2"])
        rh2 = rh * rh
        lo, hi = fma32(rh2, M["hlo"], -M["m"]), fma32(rh2, M["hhi"], M["m"])
    else:
        lo, hi = M["rh2lo"], M["rh2hi"]
    half = full & (d2 < lo)
    exact |= full & (d2 >= lo) & (d2 <= hi)
    return full, half, exact


def reference(ri, rj, Di, Dj, L, Rv, zeta, rcut, skin, berthier=True, lj_md=None):
    d = ri - rj
    d = d - np.rint(d / L) * L
    d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
    for c in range(2, ri.shape[1]):
        d2 = d2 + d[:, c] * d[:, c]
    md = (Di + Dj) * 0.5 * (1.0 - zeta * np.abs(Di - Dj)) if berthier else lj_md
    rh = rcut * md + skin
    return d2 <= Rv * Rv, (d2 <= Rv * Rv) & (d2 < rh * rh)


def adversarial_pairs(L, dims, Rv, rcut, skin, zeta, Dlo, Dhi, rng, n=60000, berthier=True, lj_md=None):
    """Pairs whose FP64 separation sits at R_v or at the pair's half-list
    radius, times (1 + eps) for eps in +-[1e-9, 1e-3] and 0, in random
    directions, with the first particle anywhere in the box (so the pair is
    often split by the periodic boundary)."""
    Di = np.exp(rng.uniform(np.log(Dlo), np.log(Dhi), n))
    Dj = np.exp(rng.uniform(np.log(Dlo), np.log(Dhi), n))
    if berthier:
        rh = rcut * (Di + Dj) * 0.5 * (1.0 - zeta * np.abs(Di - Dj)) + skin
    else:
        rh = np.full(n, rcut * lj_md + skin)
    target = np.where(rng.uniform(size=n) < 0.5, Rv, np.minimum(rh, Rv))
    eps = rng.choice([0.0, 1e-9, 1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3], n) * rng.choice([-1.0, 1.0], n)
    dirs = rng.standard_normal((n, dims))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    ri = rng.uniform(0.0, L, (n, dims))
    rj = ri + dirs * (target * (1.0 + eps))[:, None]
    rj -= np.floor(rj / L) * L
    return ri, rj, Di, Dj, eps


CASES = [
    # L, dims, diameters, zeta, potential
    (99.342, 3, (0.73, 1.62), 0.2, True),   # the N=10^6 system
    (20.0, 3, (0.73, 1.62), 0.2, True),
    (31.6, 2, (0.73, 1.62), 0.2, True),
    (1000.0, 3, (0.73, 1.62), 0.2, True),   # a huge box: coarse fixed-point units
    (8.0, 3, (0.8, 1.0), 0.5, True),        # zeta (max D - min D) = 0.1
    (25.0, 3, (0.5, 1.5), 0.5, True),       # zeta (max D - min D) = 0.5, the edge the margins allow
    (12.0, 3, (1.0, 1.0), 0.0, False),      # LJ, sigma^2 half-list quirk
    (15.0, 2, (1.0, 1.0), 0.0, False),
]


@pytest.mark.parametrize("L,dims,Drange,zeta,berthier", CASES)
def test_prefilter_decisions_equal_fp64(L, dims, Drange, zeta, berthier):
    rng = np.random.default_rng(int(L * 10) + dims)
    rcut, skin = (1.25, 0.3) if berthier else (2.5, 0.3)
    lj_md = None if berthier else 1.0
    Rv = skin + rcut * Drange[1]
    M = margins(L, Rv, dims, zeta, Drange[1], rcut, skin, lj_md=lj_md)
    ri, rj, Di, Dj, eps = adversarial_pairs(L, dims, Rv, rcut, skin, zeta, *Drange, rng, berthier=berthier,
                                            lj_md=lj_md)
    full, half, exact = prefilter(ri, rj, Di, Dj, L, M, berthier)
    rf, rhf = reference(ri, rj, Di, Dj, L, Rv, zeta, rcut, skin, berthier, lj_md)
    decided = ~exact
    assert np.array_equal(full[decided], rf[decided])
    assert np.array_equal(half[decided], rhf[decided])
    # the FP64 test is needed only close to a threshold: pairs placed 1e-4 or
    # more (relative) from the one they were aimed at never take it, unless
    # they sit near the other one too
    assert decided.any() and exact.any()
    far = np.abs(eps) >= 1e-4
    assert exact[far].mean() < 0.02


def test_planted_small_margin_is_caught():
    """Margins a hundred times too small take wrong decisions on the same
    adversarial pairs: the check above would fail."""
    L, dims, Drange, zeta = 99.342, 3, (0.73, 1.62), 0.2
    rng = np.random.default_rng(7)
    Rv = 0.3 + 1.25 * Drange[1]
    M = margins(L, Rv, dims, zeta, Drange[1], 1.25, 0.3, shrink=0.01)
    ri, rj, Di, Dj, _ = adversarial_pairs(L, dims, Rv, 1.25, 0.3, zeta, *Drange, rng)
    full, half, exact = prefilter(ri, rj, Di, Dj, L, M)
    rf, rhf = reference(ri, rj, Di, Dj, L, Rv, zeta, 1.25, 0.3)
    decided = ~exact
    assert not (np.array_equal(full[decided], rf[decided]) and np.array_equal(half[decided], rhf[decided]))


def test_fma32_is_correctly_rounded():
    """fma32 against exact rational arithmetic on random float32 triples, on
    products that cancel the addend, and on sums whose float64 rounding lands
    on a float32 midpoint (x + a b with a b = +-2^-24 (1 - 2^-36): the float64
    sum is x +- 2^-24, exactly halfway, while the exact value lies inside)."""
    from fractions import Fraction
    rng = np.random.default_rng(11)
    n = 3000
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    c = rng.standard_normal(n).astype(np.float32)
    k = slice(1000, 2000)  # near-cancellation of the product and the addend
    c[k] = (-(a[k].astype(np.float64) * b[k]) * (1.0 + rng.uniform(-1e-6, 1e-6, 1000))).astype(np.float32)
    m = slice(0, 1000)  # midpoints
    x = rng.uniform(1.0, 2.0, 1000).astype(np.float32)
    a[m] = np.float32(2.0 ** -24 * (1.0 + 2.0 ** -18))
    b[m] = np.where(rng.uniform(size=1000) < 0.5, 1.0, -1.0).astype(np.float32) * np.float32(1.0 - 2.0 ** -18)
    c[m] = x
    s64 = a[m].astype(np.float64) * b[m] + c[m]
    half = np.abs(s64 - x.astype(np.float64))
    assert np.all(half == 2.0 ** -24)  # the float64 sum sits exactly halfway
    got = fma32(a, b, c)
    for i in range(n):
        exact = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        f = np.float32(float(exact))
        cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        best = min(cands, key=lambda y: (abs(Fraction(float(y)) - exact), int(np.float32(y).view(np.uint32)) & 1))
        assert got[i] == best, (i, a[i], b[i], c[i])
    # without the midpoint correction, plain float32(float64 sum) is wrong on about half of them
    plain = s64.astype(np.float32)
    assert np.mean(plain != got[m]) > 0.2
