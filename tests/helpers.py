"""Shared helpers for the tests: system setup and independent numpy checks.

The numpy routines here are written independently of the oracle's C code
(vectorised, different loop structure) so that they can pin it.
"""
import math

import numpy as np

import oracle
from simulationcode_amd import _native as nat, synth


def berthier_consts(rc):
    """c0, c2, c4 from the C2-smoothness conditions V(rc)=V'(rc)=V''(rc)=0 of
    V(x) = x^-12 + c0 + c2 x^2 + c4 x^4 (x = r/sigma), solved as a linear
    system -- independent of the closed forms in the reference constructor."""
    A = np.array([[1.0, rc ** 2, rc ** 4],
                  [0.0, 2 * rc, 4 * rc ** 3],
                  [0.0, 2.0, 12 * rc ** 2]])
    b = -np.array([rc ** -12, -12 * rc ** -13, 156 * rc ** -14])
    return np.linalg.solve(A, b)


def np_min_image_d2(r, L):
    """All-pairs squared minimum-image distances by explicit image search."""
    dr = r[:, None, :] - r[None, :, :]
    dims = r.shape[1]
    best = np.full(dr.shape[:2], np.inf)
    shifts = np.array(np.meshgrid(*[[-1, 0, 1]] * dims, indexing="ij")).reshape(dims, -1).T
    for s in shifts:
        d = dr + s * L
        best = np.minimum(best, np.sum(d * d, axis=2))
    return best


def np_pairs(r, D, L, Rv, rc, skin, zeta, lj_sigma=None):
    """Independent enumeration of the full (d2 <= Rv^2) and half
    (also d2 < (rc*sigma_ij + skin)^2) pair sets as sorted i*N+j keys, i<j."""
    N = len(r)
    d2 = np_min_image_d2(r, L)
    iu, ju = np.triu_indices(N, 1)
    dd = d2[iu, ju]
    if lj_sigma is None:
        Di, Dj = D[iu], D[ju]
        md = (Di + Dj) * 0.5 * (1.0 - zeta * np.abs(Di - Dj))
    else:
        md = np.full(len(iu), lj_sigma * lj_sigma)
    full = dd <= Rv * Rv
    half = full & (dd < (rc * md + skin) ** 2)
    key = iu.astype(np.int64) * N + ju
    return np.sort(key[full]), np.sort(key[half]), dd, md


def make_system(kind, N, dims, seed, *, system=oracle.SYS_NEWTONIAN, kBT=0.15, dt=0.005, gamma=1.0,
                rho=None, skin=0.3, jitter=0.1, maxneighbours=200, rescale=False, relax=True):
    """Seeded system: 'lj' (eps=sigma=1, rc=2.5) or 'berthier' (rc=1.25).
    Polydisperse starts are pushed apart by a short capped steepest descent
    on the oracle so that forces are moderate."""
    rng = np.random.default_rng(seed)
    if kind == "lj":
        rho = rho or 1.0
        L = synth.box_size(N, rho, dims)
        D = np.ones(N)
        pot, rc = oracle.POT_LJ, 2.5
    else:
        rho = rho or (1.02 if dims == 3 else 1.0)
        L = synth.box_size(N, rho, dims)
        D = synth.berthier_diameters(N, rng)
        pot, rc = oracle.POT_BERTHIER, 1.25
    r = synth.lattice_positions(N, dims, L, jitter, rng)
    v = synth.maxwell_boltzmann(N, dims, kBT, 1.0, rng)
    op = oracle.make_params(dims=dims, N=N, box_size=L, force_cutoff=rc, skin=skin, D=D, potential=pot,
                            system=system, kBT=kBT, gamma=gamma, dt=dt, maxneighbours=maxneighbours,
                            rescale=rescale)
    if relax and kind == "berthier":
        r = relax_oracle(op, r, D)
    return op, r, D, v, rc


def relax_oracle(op, r, D, iters=300, step=0.02):
    r = r.copy()
    L = op.box_size
    for it in range(iters):
        lists = oracle.build_lists(op, r, D, cells=True)
        F = oracle.calculate_new_forces(op, r, D, lists)
        fmax = np.max(np.linalg.norm(F, axis=1))
        if fmax < 1.0:
            break
        dr = F * 0.005
        s = np.max(np.linalg.norm(dr, axis=1))
        if s > step:
            dr *= step / s
        r = r + dr
        r -= np.floor(r / L) * L
    return np.ascontiguousarray(r)


def sc_params_from(op, rc, precision=nat.PREC_FP64, cell_divisor=1, seed=1234, build_mode=0):
    p = nat.ScParams()
    for f, _ in nat.ScParams._fields_:
        if hasattr(op, f):
            setattr(p, f, getattr(op, f))
    p.force_cutoff = rc
    p.precision = precision
    p.cell_divisor = cell_divisor
    p.build_mode = build_mode
    p.seed = seed
    return p


def force_rel_err(Fg, Fo):
    """max_i |dF_i| / max_i |F_i| (Euclidean norms per particle)."""
    return float(np.max(np.linalg.norm(Fg - Fo, axis=1)) / max(np.max(np.linalg.norm(Fo, axis=1)), 1e-300))


def tie_band_ok(dd, thresholds, rel=1e-12):
    """No squared distance within rel of any predicate threshold."""
    for t in thresholds:
        t = np.broadcast_to(t, dd.shape)
        if np.any(np.abs(dd - t) <= rel * t):
            return False
    return True
