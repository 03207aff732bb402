"""ctypes binding of libsc (include/sc.h).

The library is built in-tree (``simulationcode_amd/_lib/libsc.so``) by
``simulationcode_amd.build``.  There is no fallback: if the library is
missing or fails to load, every entry point raises.
"""
import ctypes
import time
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsc.so")

POT_LJ = 1
POT_BERTHIER = 2
SYS_NEWTONIAN = 0
SYS_LANGEVIN = 1
SYS_BROWNIAN = 2
PREC_FP64 = 0
PREC_MIXED = 1

SC_OK = 0
SC_ERR_ARG = -1
SC_ERR_OOM = -2
SC_ERR_OVERFLOW = -3
SC_ERR_NONFINITE = -4
SC_ERR_HIP = -5
SC_ERR_STATE = -6

_ERRNAMES = {
    SC_ERR_ARG: "SC_ERR_ARG",
    SC_ERR_OOM: "SC_ERR_OOM",
    SC_ERR_OVERFLOW: "SC_ERR_OVERFLOW",
    SC_ERR_NONFINITE: "SC_ERR_NONFINITE",
    SC_ERR_HIP: "SC_ERR_HIP",
    SC_ERR_STATE: "SC_ERR_STATE",
}

# Every symbol include/sc.h declares (checked by the CPU tests).
EXPORTS = (
    "sc_create", "sc_destroy", "sc_last_error", "sc_set_state", "sc_get_state",
    "sc_update_neighbor_lists", "sc_get_neighbor_lists", "sc_get_neighbor_lists_noalloc", "sc_calculate_new_forces",
    "sc_do_time_steps", "sc_full_energy", "sc_energy_no_neigh", "sc_kinetic_energy",
    "sc_rescale_velocities", "sc_rebuild_count", "sc_minimize", "sc_minimize_until", "sc_philox4x32_10",
    "sc_last_elapsed_ms", "sc_copy_bandwidth", "sc_potential_energies", "sc_set_kernel_timing", "sc_kernel_times", "sc_neighbor_stats",
    "sc_batch_create", "sc_batch_destroy", "sc_batch_last_error", "sc_replica_device_bytes",
    "sc_batch_add", "sc_batch_remove", "sc_batch_set_graph", "sc_batch_step", "sc_batch_minimize", "sc_batch_minimize_until",
    "sc_batch_get_state", "sc_batch_set_state", "sc_batch_replica_stats", "sc_batch_info",
    "sc_batch_neighbor_stats", "sc_batch_replica_error",
    "sc_set_build_launch", "sc_batch_set_build_launch", "sc_autotune", "sc_batch_autotune",
    "sc_batch_step_async", "sc_batch_wait", "sc_batch_set_ahead",
    "sc_set_reference_positions", "sc_f2_self", "sc_steinhardt",
)


class TuneEntry(ctypes.Structure):
    """sc_tune_entry: one launch shape of the per-slot build and its measured
    device time per step."""
    _fields_ = [("block_size", ctypes.c_int32), ("lanes_per_particle", ctypes.c_int32),
                ("ms_per_step", ctypes.c_double)]


TUNE_SHAPES = 13


def _tune(fn, h, steps):
    ents = (TuneEntry * TUNE_SHAPES)()
    best = ctypes.c_int32()
    rc = fn(h, int(steps), ents, TUNE_SHAPES, ctypes.byref(best))
    table = [{"block_size": e.block_size, "lanes_per_particle": e.lanes_per_particle,
              "ms_per_step": e.ms_per_step} for e in ents]
    return rc, table, best.value


class ScError(RuntimeError):
    """A libsc call returned an error status."""

    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class NeighborOverflowError(ScError, IndexError):
    """Raised where the reference raises a BoundsError on a full neighbour row
    (src/NeighborLists.jl:86)."""


class ScParams(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32),
        ("potential", ctypes.c_int32),
        ("system", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("rescale_temperature", ctypes.c_int32),
        ("cell_divisor", ctypes.c_int32),
        ("build_mode", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("N", ctypes.c_int64),
        ("maxneighbours", ctypes.c_int64),
        ("box_size", ctypes.c_double),
        ("force_cutoff", ctypes.c_double),
        ("force_cutoff2", ctypes.c_double),
        ("skin_distance", ctypes.c_double),
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
        ("seed", ctypes.c_uint64),
    ]


_lib = None


def load(path=LIB_PATH):
    """Load libsc and declare the signatures.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"libsc is not built ({path}); run simulationcode_amd.build.build() "
            "(or __graft_entry__.build()) first -- there is no CPU fallback")
    L = ctypes.CDLL(path)
    vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    sig = {
        "sc_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.POINTER(ScParams), vp]),
        "sc_destroy": (None, [vp]),
        "sc_last_error": (ctypes.c_char_p, [vp]),
        "sc_set_state": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
        "sc_get_state": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
        "sc_update_neighbor_lists": (ctypes.c_int, [vp]),
        "sc_get_neighbor_lists": (ctypes.c_int, [vp, i64, vp, vp, vp, vp]),
        "sc_get_neighbor_lists_noalloc": (ctypes.c_int, [vp, i64, vp, vp, vp, vp]),
        "sc_calculate_new_forces": (ctypes.c_int, [vp, vp]),
        "sc_do_time_steps": (ctypes.c_int, [vp, i64, vp]),
        "sc_full_energy": (ctypes.c_int, [vp, pd]),
        "sc_energy_no_neigh": (ctypes.c_int, [vp, pd]),
        "sc_kinetic_energy": (ctypes.c_int, [vp, pd]),
        "sc_rescale_velocities": (ctypes.c_int, [vp]),
        "sc_set_reference_positions": (ctypes.c_int, [vp]),
        "sc_f2_self": (ctypes.c_int, [vp, dbl, pd]),
        "sc_steinhardt": (ctypes.c_int, [vp, ctypes.c_int32, dbl, ctypes.c_int32, pd]),
        "sc_rebuild_count": (ctypes.c_int, [vp, ctypes.POINTER(i64), i64]),
        "sc_minimize": (ctypes.c_int, [vp, i64, dbl, dbl, pd]),
        "sc_minimize_until": (ctypes.c_int, [vp, i64, dbl, dbl, dbl, ctypes.POINTER(i64), pd,
                                             ctypes.POINTER(ctypes.c_int32)]),
        "sc_philox4x32_10": (None, [vp, vp, vp]),
        "sc_last_elapsed_ms": (ctypes.c_int, [vp, pd]),
        "sc_copy_bandwidth": (ctypes.c_int, [i64, ctypes.c_int32, pd]),
        "sc_potential_energies": (ctypes.c_int, [vp, vp]),
        "sc_set_kernel_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "sc_kernel_times": (ctypes.c_int, [vp, pd, ctypes.POINTER(i64), pd, ctypes.POINTER(i64)]),
        "sc_neighbor_stats": (ctypes.c_int, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "sc_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int32]),
        "sc_batch_destroy": (None, [vp]),
        "sc_batch_last_error": (ctypes.c_char_p, [vp]),
        "sc_replica_device_bytes": (i64, [ctypes.POINTER(ScParams), vp]),
        "sc_batch_add": (ctypes.c_int, [vp, ctypes.POINTER(ScParams), vp, vp, vp, ctypes.POINTER(ctypes.c_int32)]),
        "sc_batch_remove": (ctypes.c_int, [vp, ctypes.c_int32]),
        "sc_batch_set_graph": (ctypes.c_int, [vp, ctypes.c_int32]),
        "sc_batch_set_ahead": (ctypes.c_int, [vp, ctypes.c_int32]),
        "sc_batch_step": (ctypes.c_int, [vp, i64]),
        "sc_batch_minimize": (ctypes.c_int, [vp, i64, dbl, dbl]),
        "sc_batch_minimize_until": (ctypes.c_int, [vp, i64, dbl, dbl, dbl, vp, vp]),
        "sc_batch_get_state": (ctypes.c_int, [vp, ctypes.c_int32, vp, vp, vp]),
        "sc_batch_set_state": (ctypes.c_int, [vp, ctypes.c_int32, vp, vp, vp]),
        "sc_batch_replica_stats": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                                  pd]),
        "sc_batch_replica_error": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
        "sc_batch_neighbor_stats": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "sc_batch_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(i64),
                                         ctypes.POINTER(i64), pd]),
        "sc_batch_step_async": (ctypes.c_int, [vp, i64]),
        "sc_batch_wait": (ctypes.c_int, [vp]),
        "sc_set_build_launch": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32]),
        "sc_batch_set_build_launch": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32]),
        "sc_autotune": (ctypes.c_int, [vp, i64, ctypes.POINTER(TuneEntry), ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int32)]),
        "sc_batch_autotune": (ctypes.c_int, [vp, i64, ctypes.POINTER(TuneEntry), ctypes.c_int32,
                                             ctypes.POINTER(ctypes.c_int32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    if not (a.flags["C_CONTIGUOUS"] and a.dtype == np.float64):
        raise TypeError("arrays crossing the C-ABI must be C-contiguous float64")
    return a.ctypes.data


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """GB/s of a plain device copy (read + write counted): the measured
    streaming ceiling for the HBM-bound kernels."""
    g = ctypes.c_double()
    rc = load().sc_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g))
    if rc != 0:
        raise ScError(rc, "sc_copy_bandwidth failed")
    return g.value


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().sc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


class Context:
    """One device-resident system (a NeighborListStruct + device copy of the
    ArrayStruct).  Thin RAII wrapper over sc_ctx."""

    def __init__(self, params: ScParams, D):
        self._lib = load()
        D = np.ascontiguousarray(D, dtype=np.float64)
        if D.shape != (params.N,):
            raise ValueError("D must have N entries")
        self.params = params
        self.N = int(params.N)
        self.dims = int(params.dims)
        h = ctypes.c_void_p()
        rc = self._lib.sc_create(ctypes.byref(h), ctypes.byref(params), D.ctypes.data)
        self._h = h
        if rc != SC_OK:
            msg = self._lib.sc_last_error(h).decode() if h.value else "invalid parameters"
            self.close()
            raise ScError(rc, msg)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.sc_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != SC_OK:
            msg = self._lib.sc_last_error(self._h).decode()
            if rc == SC_ERR_OVERFLOW:
                raise NeighborOverflowError(rc, msg)
            raise ScError(rc, msg)

    def _arr(self, a, per_particle=False):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        want = self.N if per_particle else self.N * self.dims
        if a.size != want:
            raise ValueError(f"array has {a.size} values, expected {want}")
        return a

    def set_state(self, r=None, v=None, F=None, D=None, disp=None):
        r, v, F, disp = (self._arr(x) for x in (r, v, F, disp))
        D = self._arr(D, per_particle=True)
        self._check(self._lib.sc_set_state(self._h, _p(r), _p(v), _p(F), _p(D), _p(disp)))

    def get_state(self, r=True, v=True, F=True, disp=False, disp2=False):
        shp = (self.N, self.dims)
        out = {}
        bufs = {}
        for name, want, shape in (("r", r, shp), ("v", v, shp), ("F", F, shp), ("disp", disp, shp),
                                  ("disp2", disp2, (self.N,))):
            bufs[name] = np.zeros(shape) if want else None
        self._check(self._lib.sc_get_state(self._h, _p(bufs["r"]), _p(bufs["v"]), _p(bufs["F"]),
                                           _p(bufs["disp"]), _p(bufs["disp2"])))
        for k, b in bufs.items():
            if b is not None:
                out[k] = b
        return out

    def update_neighbor_lists(self):
        self._check(self._lib.sc_update_neighbor_lists(self._h))

    def get_neighbor_lists(self, maxneigh=None):
        """Reference layout: counts (N,), lists as (N, maxneigh) row-per-particle
        views of the column-major maxneigh x N arrays, 1-based."""
        M = int(maxneigh or self.params.maxneighbours)
        nfull = np.zeros(self.N, dtype=np.int64)
        nhalf = np.zeros(self.N, dtype=np.int64)
        full = np.zeros((self.N, M), dtype=np.int64)
        half = np.zeros((self.N, M), dtype=np.int64)
        self._check(self._lib.sc_get_neighbor_lists(self._h, M, nfull.ctypes.data, full.ctypes.data,
                                                    nhalf.ctypes.data, half.ctypes.data))
        return nfull, full, nhalf, half

    def calculate_new_forces(self, noise=None):
        noise = self._arr(noise)
        self._check(self._lib.sc_calculate_new_forces(self._h, _p(noise)))

    def do_time_steps(self, nsteps, noise=None):
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
            if noise.size != nsteps * self.N * self.dims:
                raise ValueError("noise must hold nsteps x N x dims values")
        self._check(self._lib.sc_do_time_steps(self._h, int(nsteps), _p(noise)))

    def _scalar(self, fn):
        out = ctypes.c_double()
        self._check(fn(self._h, ctypes.byref(out)))
        return out.value

    def full_energy(self):
        return self._scalar(self._lib.sc_full_energy)

    def potential_energies(self):
        """compute_potential_energy!: E_i / 2 per particle, original order."""
        E = np.zeros(self.N)
        self._check(self._lib.sc_potential_energies(self._h, E.ctypes.data))
        return E

    def energy_no_neigh(self):
        return self._scalar(self._lib.sc_energy_no_neigh)

    def kinetic_energy(self):
        return self._scalar(self._lib.sc_kinetic_energy)

    def rescale_velocities(self):
        self._check(self._lib.sc_rescale_velocities(self._h))

    def set_reference_positions(self):
        """r_old := r (the reference configuration of f2_self)."""
        self._check(self._lib.sc_set_reference_positions(self._h))

    def f2_self(self, kx=7.2):
        out = ctypes.c_double()
        self._check(self._lib.sc_f2_self(self._h, float(kx), ctypes.byref(out)))
        return out.value

    def steinhardt(self, l, q_cutoff, averaged=False):
        out = ctypes.c_double()
        self._check(self._lib.sc_steinhardt(self._h, int(l), float(q_cutoff), 1 if averaged else 0,
                                            ctypes.byref(out)))
        return out.value

    def rebuild_count(self, set_to=-1):
        out = ctypes.c_int64()
        self._check(self._lib.sc_rebuild_count(self._h, ctypes.byref(out), int(set_to)))
        return out.value

    def minimize(self, iters, dt, max_step):
        mf = ctypes.c_double()
        self._check(self._lib.sc_minimize(self._h, int(iters), float(dt), float(max_step), ctypes.byref(mf)))
        return mf.value

    def minimize_until(self, max_iters, dt, max_step, tolerance):
        """Iterates until max(max|F|, step) <= tolerance (checked on the
        device after every move) or max_iters; returns (iterations, max |F|
        of the last iteration, converged)."""
        it, mf, cv = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int32()
        self._check(self._lib.sc_minimize_until(self._h, int(max_iters), float(dt), float(max_step),
                                                float(tolerance), ctypes.byref(it), ctypes.byref(mf),
                                                ctypes.byref(cv)))
        return it.value, mf.value, bool(cv.value)

    def last_elapsed_ms(self):
        return self._scalar(self._lib.sc_last_elapsed_ms)

    def set_kernel_timing(self, on):
        self._check(self._lib.sc_set_kernel_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        fm, dm = ctypes.c_double(), ctypes.c_double()
        fn, dn = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.sc_kernel_times(self._h, ctypes.byref(fm), ctypes.byref(fn), ctypes.byref(dm),
                                              ctypes.byref(dn)))
        return {"force_ms": fm.value, "force_launches": fn.value, "drift_ms": dm.value,
                "drift_launches": dn.value}

    def neighbor_stats(self):
        t, m = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.sc_neighbor_stats(self._h, ctypes.byref(t), ctypes.byref(m)))
        return t.value, m.value

    def set_build_launch(self, block_size, lanes_per_particle):
        self._check(self._lib.sc_set_build_launch(self._h, int(block_size), int(lanes_per_particle)))

    def autotune(self, steps=100):
        """Time every build launch shape over `steps` real steps each, keep
        the fastest; returns (table, index of the chosen shape)."""
        rc, table, best = _tune(self._lib.sc_autotune, self._h, steps)
        self._check(rc)
        return table, best


def replica_device_bytes(params: ScParams, D):
    D = np.ascontiguousarray(D, dtype=np.float64)
    return int(load().sc_replica_device_bytes(ctypes.byref(params), D.ctypes.data))


class Batch:
    """Replicas advanced by shared launches (sc_batch_*)."""

    def __init__(self, max_replicas=256, graph=False):
        self._lib = load()
        self.max_replicas = int(max_replicas)
        h = ctypes.c_void_p()
        rc = self._lib.sc_batch_create(ctypes.byref(h), int(max_replicas))
        self._h = h
        self._shape = {}
        if rc != SC_OK:
            msg = self._lib.sc_batch_last_error(h).decode() if h.value else "invalid arguments"
            self.close()
            raise ScError(rc, msg)
        if graph:
            self.set_graph(True)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.sc_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != SC_OK:
            msg = self._lib.sc_batch_last_error(self._h).decode()
            if rc == SC_ERR_OVERFLOW:
                raise NeighborOverflowError(rc, msg)
            raise ScError(rc, msg)

    def add(self, params: ScParams, D, r, v=None):
        D = np.ascontiguousarray(D, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        if D.shape != (params.N,) or r.size != params.N * params.dims:
            raise ValueError("D / r shapes do not match params")
        if v is not None:
            v = np.ascontiguousarray(v, dtype=np.float64)
            if v.size != r.size:
                raise ValueError("v shape does not match r")
        rid = ctypes.c_int32()
        self._check(self._lib.sc_batch_add(self._h, ctypes.byref(params), D.ctypes.data, r.ctypes.data,
                                           None if v is None else v.ctypes.data, ctypes.byref(rid)))
        self._shape[rid.value] = (int(params.N), int(params.dims))
        return rid.value

    def remove(self, rid):
        self._check(self._lib.sc_batch_remove(self._h, int(rid)))
        self._shape.pop(rid, None)

    def set_graph(self, on):
        self._check(self._lib.sc_batch_set_graph(self._h, 1 if on else 0))

    def set_ahead(self, chunk):
        """Steps launched ahead per chunk (True: the default 8), or False / 0
        for eager per-step launches; the trajectories are the same."""
        chunk = 8 if chunk is True else int(chunk)
        self._check(self._lib.sc_batch_set_ahead(self._h, chunk))

    def step(self, nsteps):
        self._check(self._lib.sc_batch_step(self._h, int(nsteps)))

    def minimize(self, iters, dt, max_step):
        self._check(self._lib.sc_batch_minimize(self._h, int(iters), float(dt), float(max_step)))

    def minimize_until(self, max_iters, dt, max_step, tolerance):
        """Every replica descends until its own max(max|F|, step) <= tolerance
        or max_iters; returns {replica id: (iterations, converged)}."""
        it = np.full(self.max_replicas, -1, dtype=np.int64)
        cv = np.zeros(self.max_replicas, dtype=np.int32)
        self._check(self._lib.sc_batch_minimize_until(self._h, int(max_iters), float(dt), float(max_step),
                                                      float(tolerance), it.ctypes.data, cv.ctypes.data))
        return {rid: (int(it[rid]), bool(cv[rid])) for rid in self._shape if it[rid] >= 0}

    def get_state(self, rid):
        n, d = self._shape[rid]
        r, v, F = np.zeros((n, d)), np.zeros((n, d)), np.zeros((n, d))
        self._check(self._lib.sc_batch_get_state(self._h, int(rid), r.ctypes.data, v.ctypes.data, F.ctypes.data))
        return {"r": r, "v": v, "F": F}

    def set_state(self, rid, r=None, v=None, F=None):
        n, d = self._shape[rid]
        arrs = []
        for a in (r, v, F):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != n * d:
                raise ValueError("array shape does not match the replica")
            arrs.append(a)
        self._check(self._lib.sc_batch_set_state(self._h, int(rid), *[None if a is None else a.ctypes.data
                                                                     for a in arrs]))

    def stats(self, rid, energy=False):
        st, rb = ctypes.c_int64(), ctypes.c_int64()
        e = ctypes.c_double()
        self._check(self._lib.sc_batch_replica_stats(self._h, int(rid), ctypes.byref(st), ctypes.byref(rb),
                                                     ctypes.byref(e) if energy else None))
        out = {"steps": st.value, "rebuilds": rb.value}
        if energy:
            out["epot"] = e.value
        return out

    def replica_error(self, rid):
        c = ctypes.c_int32()
        self._check(self._lib.sc_batch_replica_error(self._h, int(rid), ctypes.byref(c)))
        return c.value

    def neighbor_stats(self, rid):
        t, m = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.sc_batch_neighbor_stats(self._h, int(rid), ctypes.byref(t), ctypes.byref(m)))
        return t.value, m.value

    def info(self):
        a = ctypes.c_int32()
        npart, by = ctypes.c_int64(), ctypes.c_int64()
        ms = ctypes.c_double()
        self._check(self._lib.sc_batch_info(self._h, ctypes.byref(a), ctypes.byref(npart), ctypes.byref(by),
                                            ctypes.byref(ms)))
        return {"active": a.value, "particles": npart.value, "device_bytes": by.value, "last_ms": ms.value}

    def set_build_launch(self, block_size, lanes_per_particle):
        self._check(self._lib.sc_batch_set_build_launch(self._h, int(block_size), int(lanes_per_particle)))

    def step_async(self, nsteps):
        """Enqueue nsteps steps on this batch's stream and return at once."""
        self._check(self._lib.sc_batch_step_async(self._h, int(nsteps)))

    def wait(self):
        """Wait for step_async's steps (updates timing, counters, errors)."""
        self._check(self._lib.sc_batch_wait(self._h))

    def autotune(self, steps=100):
        """Time every build launch shape over `steps` real steps of the batch
        each, keep the fastest; returns (table, index of the chosen shape)."""
        rc, table, best = _tune(self._lib.sc_batch_autotune, self._h, steps)
        self._check(rc)
        return table, best


class BatchGroup:
    """Replicas spread over `streams` batches, each with its own HIP stream,
    stepped concurrently: the kernels of one batch fill the gaps left by the
    latency-bound launches of the others.  A replica goes to the batch with
    the fewest particles; its trajectory does not depend on the grouping."""

    def __init__(self, max_replicas=256, streams=2, graph=True):
        self.batches = [Batch(max_replicas, graph=graph) for _ in range(int(streams))]
        self.load = [0] * len(self.batches)
        self.where = {}
        self._next = 0
        self.last_ms = 0.0

    def add(self, params, D, r, v=None):
        k = min(range(len(self.batches)), key=lambda q: self.load[q])
        rid = self.batches[k].add(params, D, r, v)
        gid = self._next
        self._next += 1
        self.where[gid] = (k, rid)
        self.load[k] += int(params.N)
        return gid

    def remove(self, gid):
        k, rid = self.where.pop(gid)
        n = self.batches[k].info()["particles"]
        self.batches[k].remove(rid)
        self.load[k] -= n - self.batches[k].info()["particles"]

    def _each(self, name, *a):
        for b in self.batches:
            getattr(b, name)(*a)

    def set_build_launch(self, block_size, lanes):
        self._each("set_build_launch", block_size, lanes)

    def set_ahead(self, chunk):
        self._each("set_ahead", chunk)

    def minimize(self, iters, dt, max_step):
        self._each("minimize", iters, dt, max_step)

    def step(self, nsteps):
        """Steps every batch concurrently; last_ms is the wall time of the
        slowest (events on each stream bracket its own work)."""
        t0 = time.perf_counter()
        live = [b for b in self.batches if b.info()["active"]]
        for b in live:
            b.step_async(nsteps)
        for b in live:
            b.wait()
        self.last_ms = (time.perf_counter() - t0) * 1e3

    def get_state(self, gid):
        k, rid = self.where[gid]
        return self.batches[k].get_state(rid)

    def stats(self, gid, energy=False):
        k, rid = self.where[gid]
        return self.batches[k].stats(rid, energy)

    def replica_error(self, gid):
        k, rid = self.where[gid]
        return self.batches[k].replica_error(rid)

    def neighbor_stats(self, gid):
        k, rid = self.where[gid]
        return self.batches[k].neighbor_stats(rid)

    def autotune(self, steps=100):
        """Tunes each batch's build launch shape on its own (one after the
        other); returns the first batch's (table, chosen index)."""
        res = [b.autotune(steps) for b in self.batches if b.info()["active"]]
        return res[0]

    def info(self):
        infos = [b.info() for b in self.batches]
        return {"active": sum(i["active"] for i in infos), "particles": sum(i["particles"] for i in infos),
                "device_bytes": sum(i["device_bytes"] for i in infos), "last_ms": self.last_ms}

    def close(self):
        self._each("close")

