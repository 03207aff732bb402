"""TEST INFRASTRUCTURE ONLY -- the oracle.

ctypes wrapper over ``oracle/liboracle_sc.so``, a single-threaded FP64 C
restatement of SimulationCode.jl's MD hot path (see ``sc_oracle.c`` for the
per-function reference citations).  Only ``tests/``, ``__graft_entry__.smoke``
and ``bench.py``'s ``cpu_baseline`` leg import this module, and only as the
checker / CPU baseline; the product never routes through it.

Pinning (checks that hold the oracle itself to the reference semantics,
all in ``tests/test_oracle.py``):
  * potential constants: closed forms of ``src/DataStructs.jl:15-18, 53-58``
    and C2 smoothness V = V' = V'' = 0 at the cutoff;
  * pair forces: analytic two-particle forces from -dV/dr written
    independently, central finite differences of the brute-force energy,
    Newton's third law;
  * neighbour lists: an independent numpy all-pairs enumeration, and the
    cell-list build against the O(N^2) build;
  * energy: list energy == O(N^2) energy (the reference's own runtime
    invariant, ``src/MD.jl:612, 623``);
  * integrators: velocity-Verlet time reversibility, NVE conservation, zero
    centre-of-mass velocity, the free-particle Langevin velocity recurrence,
    Brownian free diffusion and its zero-noise steepest-descent limit;
  * golden fixtures under ``tests/golden`` written by the oracle and read
    back.
Parity unpinned (no reference-held vector exists and Julia is absent):
CellListMap's in-row pair order and half-list orientation, LoopVectorization's
summation order, and Julia's RNG streams (noise is injected instead).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sc.so")

POT_LJ = 1
POT_BERTHIER = 2
SYS_NEWTONIAN = 0
SYS_LANGEVIN = 1
SYS_BROWNIAN = 2


class OrParams(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32),
        ("potential", ctypes.c_int32),
        ("system", ctypes.c_int32),
        ("rescale_temperature", ctypes.c_int32),
        ("N", ctypes.c_int64),
        ("maxneighbours", ctypes.c_int64),
        ("box_size", ctypes.c_double),
        ("force_cutoff2", ctypes.c_double),
        ("skin_distance", ctypes.c_double),
        ("verlet_sphere_size", ctypes.c_double),
        ("lj_eps", ctypes.c_double),
        ("lj_sigma", ctypes.c_double),
        ("lj_Uc", ctypes.c_double),
        ("b_c0", ctypes.c_double),
        ("b_c2", ctypes.c_double),
        ("b_c4", ctypes.c_double),
        ("b_zeta", ctypes.c_double),
        ("kBT", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("m", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(OrParams)
        dp = ctypes.c_void_p
        L.or_lj_Uc.restype = ctypes.c_double
        L.or_lj_Uc.argtypes = [ctypes.c_double] * 3
        L.or_berthier_constants.restype = None
        L.or_berthier_constants.argtypes = [ctypes.c_double, dp]
        L.or_squared_distance.restype = ctypes.c_double
        L.or_squared_distance.argtypes = [dp, dp, ctypes.c_int, ctypes.c_double]
        for name in ("or_build_lists_brute", "or_build_lists_cells"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [P, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_calculate_new_forces.restype = None
        L.or_calculate_new_forces.argtypes = [P, dp, dp, dp, dp, dp, dp, dp]
        L.or_full_energy.restype = ctypes.c_double
        L.or_full_energy.argtypes = [P, dp, dp, dp, dp]
        L.or_energy_no_neigh.restype = ctypes.c_double
        L.or_energy_no_neigh.argtypes = [P, dp, dp]
        L.or_kinetic_energy.restype = ctypes.c_double
        L.or_kinetic_energy.argtypes = [P, dp]
        L.or_rescale_velocities.restype = None
        L.or_rescale_velocities.argtypes = [P, dp]
        L.or_do_time_steps.restype = ctypes.c_int
        L.or_do_time_steps.argtypes = [P, ctypes.c_int64, ctypes.c_int] + [dp] * 11 + [dp]
        L.or_update_neighbor_lists.restype = ctypes.c_int
        L.or_update_neighbor_lists.argtypes = [P, ctypes.c_int] + [dp] * 8 + [dp]
        L.or_f2_self.argtypes = [P, dp, dp, ctypes.c_double]
        L.or_f2_self.restype = ctypes.c_double
        L.or_ylm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
        for f in (L.or_ql, L.or_ql_averaged):
            f.argtypes = [P, dp, dp, dp, ctypes.c_int, ctypes.c_double]
            f.restype = ctypes.c_double
        L.or_potential_energies.argtypes = [P, dp, dp, dp, dp, dp]
        L.or_find_inherent_structure.restype = ctypes.c_int
        L.or_find_inherent_structure.argtypes = ([P, ctypes.c_int] + [dp] * 10 +
                                                 [ctypes.c_double, ctypes.c_int64, ctypes.c_double,
                                                  ctypes.c_double, dp, dp, dp])
        L.or_baseline_steps.restype = ctypes.c_int
        L.or_baseline_steps.argtypes = [P, ctypes.c_int64, ctypes.c_int, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def lj_Uc(eps, sigma, rc):
    return lib().or_lj_Uc(eps, sigma, rc)


def berthier_constants(rc):
    out = np.zeros(3)
    lib().or_berthier_constants(rc, _ptr(out))
    return tuple(out)


def squared_distance(ri, rj, L):
    ri = np.ascontiguousarray(ri, dtype=np.float64)
    rj = np.ascontiguousarray(rj, dtype=np.float64)
    return lib().or_squared_distance(_ptr(ri), _ptr(rj), len(ri), L)


def make_params(*, dims, N, box_size, force_cutoff, skin, D, potential, system,
                maxneighbours=200, eps=1.0, sigma=1.0, zeta=0.2, kBT=1.0, gamma=1.0,
                dt=0.01, m=1.0, rescale=False):
    """Fill OrParams the way ParameterStruct/initialize_neighbor_struct do
    (src/DataStructs.jl:138-153, 349-354)."""
    p = OrParams()
    p.dims = dims
    p.potential = potential
    p.system = system
    p.rescale_temperature = 1 if rescale else 0
    p.N = N
    p.maxneighbours = maxneighbours
    p.box_size = box_size
    p.force_cutoff2 = force_cutoff * force_cutoff
    p.skin_distance = skin
    p.verlet_sphere_size = skin + force_cutoff * float(np.max(D))
    p.lj_eps = eps
    p.lj_sigma = sigma
    p.lj_Uc = lj_Uc(eps, sigma, force_cutoff) if potential == POT_LJ else 0.0
    c = berthier_constants(force_cutoff)
    p.b_c0, p.b_c2, p.b_c4 = c
    p.b_zeta = zeta
    p.kBT, p.gamma, p.dt, p.m = kBT, gamma, dt, m
    return p


class Lists:
    """Reference-layout lists (0-based, row per particle)."""

    def __init__(self, N, M):
        self.nfull = np.zeros(N, dtype=np.int64)
        self.full = np.zeros((N, M), dtype=np.int64)
        self.nhalf = np.zeros(N, dtype=np.int64)
        self.half = np.zeros((N, M), dtype=np.int64)

    def full_pairs(self):
        return _pairs(self.nfull, self.full, both=True)

    def half_pairs(self):
        return _pairs(self.nhalf, self.half, both=False)


def _pairs(n, lst, both):
    rows = np.repeat(np.arange(len(n)), n)
    cols = np.concatenate([lst[i, : n[i]] for i in range(len(n))]) if len(n) else np.zeros(0, np.int64)
    a = np.minimum(rows, cols)
    b = np.maximum(rows, cols)
    key = a.astype(np.int64) * len(n) + b
    if both:
        key = np.unique(key)
    return np.sort(key)


def build_lists(p, r, D, cells=False, disp=None, disp2=None):
    L_ = Lists(p.N, p.maxneighbours)
    f = lib().or_build_lists_cells if cells else lib().or_build_lists_brute
    rc = f(ctypes.byref(p), _ptr(r), _ptr(D), _ptr(L_.nfull), _ptr(L_.full),
           _ptr(L_.nhalf), _ptr(L_.half), _ptr(disp), _ptr(disp2))
    if rc != 0:
        raise RuntimeError(f"oracle list build failed: {rc}")
    return L_


def calculate_new_forces(p, r, D, lists, v=None, noise=None):
    F = np.zeros_like(r)
    lib().or_calculate_new_forces(ctypes.byref(p), _ptr(r), _ptr(D), _ptr(lists.nhalf),
                                  _ptr(lists.half), _ptr(v), _ptr(noise), _ptr(F))
    return F


def potential_energies(p, r, D, lists):
    """compute_potential_energy!: E_i / 2 per particle."""
    E = np.zeros(p.N)
    lib().or_potential_energies(ctypes.byref(p), _ptr(r), _ptr(D), _ptr(lists.nfull), _ptr(lists.full), _ptr(E))
    return E


def full_energy(p, r, D, lists):
    return lib().or_full_energy(ctypes.byref(p), _ptr(r), _ptr(D), _ptr(lists.nfull), _ptr(lists.full))


def energy_no_neigh(p, r, D):
    return lib().or_energy_no_neigh(ctypes.byref(p), _ptr(r), _ptr(D))


def kinetic_energy(p, v):
    return lib().or_kinetic_energy(ctypes.byref(p), _ptr(v))


def rescale_velocities(p, v):
    lib().or_rescale_velocities(ctypes.byref(p), _ptr(v))


def baseline_steps(p, nsteps, nthreads, r, v, F, D):
    """All-core CPU timing baseline (not a checker): nsteps Newtonian
    velocity-Verlet steps on nthreads OpenMP threads, lists rebuilt on entry.
    r, v, F (N x dims FP64) are advanced in place; returns the rebuilds."""
    for a in (r, v, F):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    D = np.ascontiguousarray(D, dtype=np.float64)
    rebuilds = np.zeros(1, dtype=np.int64)
    rc = lib().or_baseline_steps(ctypes.byref(p), int(nsteps), int(nthreads), _ptr(r), _ptr(v), _ptr(F),
                                 _ptr(D), _ptr(rebuilds))
    if rc != 0:
        raise RuntimeError(f"baseline steps failed: {rc}")
    return int(rebuilds[0])


class State:
    """Mutable MD state held in numpy (AoS N x dims FP64, as ArrayStruct)."""

    def __init__(self, p, r, v, D, F=None):
        self.p = p
        self.r = np.ascontiguousarray(r, dtype=np.float64).copy()
        self.v = np.ascontiguousarray(v, dtype=np.float64).copy()
        self.D = np.ascontiguousarray(D, dtype=np.float64).copy()
        self.F = np.zeros_like(self.r) if F is None else np.ascontiguousarray(F, dtype=np.float64).copy()
        self.disp = np.zeros_like(self.r)
        self.disp2 = np.zeros(p.N)
        self.lists = Lists(p.N, p.maxneighbours)
        self.rebuilds = np.zeros(1, dtype=np.int64)

    def update_neighbor_lists(self, cells=True):
        rc = lib().or_update_neighbor_lists(
            ctypes.byref(self.p), 1 if cells else 0, _ptr(self.r), _ptr(self.D),
            _ptr(self.lists.nfull), _ptr(self.lists.full), _ptr(self.lists.nhalf),
            _ptr(self.lists.half), _ptr(self.disp), _ptr(self.disp2), _ptr(self.rebuilds))
        if rc != 0:
            raise RuntimeError(f"oracle rebuild failed: {rc}")

    def do_time_steps(self, nsteps, noise=None, cells=True):
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
            assert noise.size >= nsteps * self.r.size
        rc = lib().or_do_time_steps(
            ctypes.byref(self.p), nsteps, 1 if cells else 0, _ptr(self.r), _ptr(self.v),
            _ptr(self.F), _ptr(self.D), _ptr(self.disp), _ptr(self.disp2),
            _ptr(self.lists.nfull), _ptr(self.lists.full), _ptr(self.lists.nhalf),
            _ptr(self.lists.half), _ptr(noise), _ptr(self.rebuilds))
        if rc != 0:
            raise RuntimeError(f"oracle time steps failed: {rc}")

    def calculate_new_forces(self, noise=None):
        self.F = calculate_new_forces(self.p, self.r, self.D, self.lists, self.v, noise)
        return self.F

    def find_inherent_structure(self, tolerance=1e-6, max_iterations=10 ** 6, max_step=0.001, dt=0.005,
                                cells=True):
        """src/InherentStructure.jl:4-56 on r in place (v, F, lists and
        displacements mutated too).  Returns (iterations, max |F| of the last
        iteration, converged); not converged after max_iterations iterations
        is reported, not raised."""
        it = np.zeros(1, dtype=np.int64)
        fm = np.zeros(1)
        rc = lib().or_find_inherent_structure(
            ctypes.byref(self.p), 1 if cells else 0, _ptr(self.r), _ptr(self.D), _ptr(self.F), _ptr(self.v),
            _ptr(self.disp), _ptr(self.disp2), _ptr(self.lists.nfull), _ptr(self.lists.full),
            _ptr(self.lists.nhalf), _ptr(self.lists.half), float(tolerance), int(max_iterations),
            float(max_step), float(dt), _ptr(it), _ptr(self.rebuilds), _ptr(fm))
        if rc not in (0, -3):
            raise RuntimeError(f"oracle minimiser failed: {rc}")
        return int(it[0]), float(fm[0]), rc == 0

    def full_energy(self):
        return full_energy(self.p, self.r, self.D, self.lists)


def f2_self(p, r_old, r, kx=7.2):
    """calculate_F2_self! (src/AuxiliaryFunctions.jl:968-987)."""
    r_old = np.ascontiguousarray(r_old, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return lib().or_f2_self(ctypes.byref(p), _ptr(r_old), _ptr(r), kx)


def ylm(l, m, theta, phi):
    """Y_l^m(theta, phi) as the oracle evaluates it (complex)."""
    re = np.zeros(1)
    im = np.zeros(1)
    lib().or_ylm(int(l), int(m), float(theta), float(phi), _ptr(re), _ptr(im))
    return complex(re[0], im[0])


def steinhardt(p, r, lists, l, q_cutoff, averaged=False):
    """calculate_ql / calculate_ql_averaged (src/AuxiliaryFunctions.jl:687-902)
    over the oracle's full lists."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    f = lib().or_ql_averaged if averaged else lib().or_ql
    return f(ctypes.byref(p), _ptr(r), _ptr(lists.nfull), _ptr(lists.full), int(l), float(q_cutoff))
