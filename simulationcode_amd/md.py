"""Host-side mirror of SimulationCode.jl's MD API over libsc.

Same names, argument meaning and order as the reference (Julia's trailing
``!`` dropped: ``update_neighbor_lists!`` -> ``update_neighbor_lists``).
State lives in numpy arrays with the reference's host layout (N x dims,
C-contiguous == Julia's Vector{SVector{dims,Float64}}); the device copy is
owned by the NeighborListStruct's libsc context and kept resident by
``perform_molecular_dynamics`` between host-visible events.

Reference map:
  potential / system types ............ src/DataStructs.jl:10-87
  ParameterStruct / ArrayStruct ........ src/DataStructs.jl:138-201
  OutputStruct ......................... src/DataStructs.jl:291-315
  initialize_neighbor_struct ........... src/DataStructs.jl:349-378
  update_neighbor_lists ................ src/NeighborLists.jl:24-53
  calculate_new_forces ................. src/MD.jl:262-316
  do_time_step ......................... src/MD.jl:429-554
  rescale_velocities ................... src/MD.jl:391-402
  run_short_MD_equilibration ........... src/MD.jl:579-586
  perform_molecular_dynamics ........... src/MD.jl:680-719
  calculate_full_energy ................ src/AuxiliaryFunctions.jl:22-34
  find_kinetic_energy .................. src/AuxiliaryFunctions.jl:57-65
  calculate_full_energy_no_neigh ....... src/AuxiliaryFunctions.jl:86-110
  calculate_ql / calculate_ql_averaged . src/AuxiliaryFunctions.jl:687-711, 832-902
  logdata .............................. src/AuxiliaryFunctions.jl:637-646
  calculate_F2_self .................... src/AuxiliaryFunctions.jl:968-987
  print_log_data (in the MD loop) ...... src/MD.jl:610-650, crystallisation guard :711-713
  create_when_to_save_array ............ src/Dump.jl:18-31
  save_data / prepare_savefile ......... src/Dump.jl:60-88, 121-156
  compute_potential_energy ............. src/AuxiliaryFunctions.jl:186-196
  read_restart ......................... src/Restart.jl:1-104
  find_inherent_structure .............. src/InherentStructure.jl:5-87
"""
import json
import math
import os
import time
import unicodedata

import numpy as np

from . import _native as nat


# ----------------------------------------------------------------- types

class InteractionPotential:
    pass


class LJ(InteractionPotential):
    """src/DataStructs.jl:10-19."""

    def __init__(self, ϵ, σ, rc):
        self.ϵ = float(ϵ)
        self.σ = float(σ)
        self.rc = float(rc)
        q = self.σ / self.rc
        self.Uc = 4.0 * self.ϵ * (q ** 12 - q ** 6)


class Berthier(InteractionPotential):
    """src/DataStructs.jl:45-60 (n = 12 only, as asserted there)."""

    def __init__(self, *, n=12, rc=1.25, σ_ratio=2.219, ζ=0.2):
        if n != 12:
            raise AssertionError("n == 12")
        self.c0 = -(1 / 8) * (8 + 6 * n + n ** 2) * rc ** -n
        self.c2 = 1 / 4 * n * (4 + n) * rc ** (-2 - n)
        self.c4 = -(1 / 8) * n * (2 + n) * rc ** (-4 - n)
        self.ζ = float(ζ)
        self.σ_ratio = float(σ_ratio)
        self.n = n
        self.rc = float(rc)


class System:
    pass


class Brownian(System):
    """src/DataStructs.jl:64-69."""

    def __init__(self, kBT, γ, Δt, dims):
        self.kBT, self.γ, self.Δt, self.dims = float(kBT), float(γ), float(Δt), int(dims)


class Newtonian(System):
    """src/DataStructs.jl:71-77."""

    def __init__(self, kBT, Δt, m, rescale_temperature, dims):
        self.kBT, self.Δt, self.m = float(kBT), float(Δt), float(m)
        self.rescale_temperature = bool(rescale_temperature)
        self.dims = int(dims)


class Langevin(System):
    """src/DataStructs.jl:79-87 (rescale_temperature is always false)."""

    def __init__(self, kBT, γ, Δt, m, dims):
        self.kBT, self.γ, self.Δt, self.m = float(kBT), float(γ), float(Δt), float(m)
        self.rescale_temperature = False
        self.dims = int(dims)


class DumpInfo:
    """src/DataStructs.jl:90-99."""

    def __init__(self, save, filename, when_to_save_array, save_r, save_v, save_F, save_D, save_Epot):
        self.save = bool(save)
        self.filename = filename
        self.when_to_save_array = when_to_save_array
        self.save_r, self.save_v, self.save_F = save_r, save_v, save_F
        self.save_D, self.save_Epot = save_D, save_Epot


class ParameterStruct:
    """src/DataStructs.jl:138-153 (positional, same order)."""

    def __init__(self, N_MD_equilibration_steps, random_seed, ρ, N, box_size, N_steps, swap_probability,
                 max_MC_displacement, force_cutoff2, q_cutoff, system, interaction_potential, dump_info,
                 callback):
        self.N_MD_equilibration_steps = int(N_MD_equilibration_steps)
        self.random_seed = int(random_seed)
        self.ρ = float(ρ)
        self.N = int(N)
        self.box_size = float(box_size)
        self.N_steps = int(N_steps)
        self.swap_probability = swap_probability
        self.max_MC_displacement = max_MC_displacement
        self.force_cutoff2 = float(force_cutoff2)
        self.q_cutoff = float(q_cutoff)
        self.system = system
        self.interaction_potential = interaction_potential
        self.dump_info = dump_info
        self.callback = callback


class ArrayStruct:
    """src/DataStructs.jl:183-201: AoS N x dims FP64 arrays."""

    def __init__(self, N, dims):
        z = lambda: np.zeros((N, dims))
        self.r_array = z()
        self.r_old_array = z()
        self.D_array = np.zeros(N)
        self.v_array = z()
        self.F_array = z()
        self.random_displacement = z()
        self.potential_energy_array = np.zeros(N)
        self.dims = dims


class OutputStruct:
    """src/DataStructs.jl:291-315."""

    def __init__(self):
        self.potential_energy = 0.0
        self.kinetic_energy = 0.0
        self.steps_done = 0
        self.N_translations_accepted = 0
        self.N_swaps_accepted = 0
        self.N_translations_total = 0
        self.N_swaps_total = 0
        self.N_neighbor_list_rebuilds = 0
        self.q4 = 0.0
        self.q6 = 0.0
        self.F2s = 0.0
        # the last log step's calculate_full_energy_no_neigh (the reference
        # prints it as E_pot_test without storing it, src/MD.jl:623, 628)
        self.E_pot_test = None


class NeighborListStruct:
    """src/DataStructs.jl:243-257.  The lists themselves live on the device;
    the reference-layout arrays are materialised on access."""

    def __init__(self, skin_distance, box_size, force_cutoff, N, D_array, dims, maxneighbours,
                 precision, cell_divisor):
        σ_max = float(np.max(D_array))
        self.skin_distance = float(skin_distance)
        self.force_cutoff = float(force_cutoff)
        self.verlet_sphere_size = self.skin_distance + self.force_cutoff * σ_max
        self.N_cells = int(math.floor(box_size / self.verlet_sphere_size))
        self.cell_size = box_size / self.N_cells if self.N_cells > 0 else box_size
        self.maxneighbours = int(maxneighbours)
        self.displacement_array = np.zeros((N, dims))
        self.displacement_squared_array = np.zeros(N)
        self.precision = precision
        self.cell_divisor = int(cell_divisor)
        self.N = N
        self.dims = dims
        self.ctx = None
        self._key = None

    # reference-layout lists (column-major maxneigh x N, 1-based)
    def _lists(self):
        if self.ctx is None:
            raise RuntimeError("neighbour lists have not been built")
        nfull, full, nhalf, half = self.ctx.get_neighbor_lists(self.maxneighbours)
        return nfull, full.T, nhalf, half.T

    @property
    def neighbor_numbers_full(self):
        return self._lists()[0]

    @property
    def neighbor_list_full(self):
        return self._lists()[1]

    @property
    def neighbor_numbers_half(self):
        return self._lists()[2]

    @property
    def neighbor_list_half(self):
        return self._lists()[3]


def initialize_neighbor_struct(skin_distance, box_size, force_cutoff, N, D_array, dims, *, maxneighbours=200,
                               precision="fp64", cell_divisor=1):
    """src/DataStructs.jl:349-378.  Extra keywords (no reference
    counterpart): ``precision`` ("fp64" | "mixed") and ``cell_divisor``
    (cells of side >= R_v / cell_divisor)."""
    if dims not in (2, 3):
        raise RuntimeError("To be implemented")
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    return NeighborListStruct(skin_distance, box_size, force_cutoff, N, D_array, dims, maxneighbours,
                              precision, cell_divisor)


# ------------------------------------------------------- device binding

def _sys_code(system):
    if isinstance(system, Newtonian):
        return nat.SYS_NEWTONIAN
    if isinstance(system, Langevin):
        return nat.SYS_LANGEVIN
    if isinstance(system, Brownian):
        return nat.SYS_BROWNIAN
    raise TypeError("unsupported system")


def make_params(parameters, neighborlist, seed=None):
    """sc_params from the reference structs."""
    U = parameters.interaction_potential
    S = parameters.system
    p = nat.ScParams()
    p.dims = S.dims
    if isinstance(U, LJ):
        p.potential = nat.POT_LJ
        p.lj_eps, p.lj_sigma, p.lj_Uc = U.ϵ, U.σ, U.Uc
    elif isinstance(U, Berthier):
        p.potential = nat.POT_BERTHIER
        p.b_c0, p.b_c2, p.b_c4, p.b_zeta = U.c0, U.c2, U.c4, U.ζ
    else:
        raise TypeError("only LJ and Berthier potentials have MD force kernels")
    p.system = _sys_code(S)
    p.precision = nat.PREC_MIXED if neighborlist.precision == "mixed" else nat.PREC_FP64
    p.rescale_temperature = 1 if getattr(S, "rescale_temperature", False) else 0
    p.cell_divisor = neighborlist.cell_divisor
    p.N = parameters.N
    p.maxneighbours = neighborlist.maxneighbours
    p.box_size = parameters.box_size
    p.force_cutoff = neighborlist.force_cutoff
    p.force_cutoff2 = parameters.force_cutoff2
    p.skin_distance = neighborlist.skin_distance
    p.kBT = S.kBT
    p.gamma = getattr(S, "γ", 0.0)
    p.dt = S.Δt
    p.m = getattr(S, "m", 1.0)
    p.seed = (parameters.random_seed if seed is None else seed) & 0xFFFFFFFFFFFFFFFF
    return p


def _bind(arrays, parameters, neighborlist):
    key = (id(parameters.system), id(parameters.interaction_potential), parameters.N, parameters.box_size,
           type(parameters.system).__name__, getattr(parameters.system, "Δt", None),
           getattr(parameters.system, "kBT", None), getattr(parameters.system, "rescale_temperature", None))
    if neighborlist.ctx is None or neighborlist._key != key:
        if neighborlist.ctx is not None:
            neighborlist.ctx.close()
        neighborlist.ctx = nat.Context(make_params(parameters, neighborlist), arrays.D_array)
        neighborlist._key = key
    return neighborlist.ctx


def _upload(arrays, neighborlist, ctx):
    ctx.set_state(r=arrays.r_array, v=arrays.v_array, F=arrays.F_array, D=arrays.D_array,
                  disp=neighborlist.displacement_array)


def _download(arrays, neighborlist, ctx):
    st = ctx.get_state(r=True, v=True, F=True, disp=True, disp2=True)
    arrays.r_array[...] = st["r"]
    arrays.v_array[...] = st["v"]
    arrays.F_array[...] = st["F"]
    neighborlist.displacement_array[...] = st["disp"]
    neighborlist.displacement_squared_array[...] = st["disp2"]


def _sync_rebuilds(output, ctx, before):
    output.N_neighbor_list_rebuilds += ctx.rebuild_count() - before


# ------------------------------------------------------------ hot path

_rng = np.random.default_rng()


def seed(s):
    """Random.seed! for the host noise stream."""
    global _rng
    _rng = np.random.default_rng(s)


def update_neighbor_lists(arrays, parameters, output, neighborlist):
    """src/NeighborLists.jl:24-53."""
    ctx = _bind(arrays, parameters, neighborlist)
    _upload(arrays, neighborlist, ctx)
    before = ctx.rebuild_count()
    ctx.update_neighbor_lists()
    _sync_rebuilds(output, ctx, before)
    neighborlist.displacement_array[...] = 0.0
    neighborlist.displacement_squared_array[...] = 0.0


def calculate_new_forces(arrays, parameters, neighborlist, noise=None):
    """src/MD.jl:262-316 (+ add_new_random_force!, :30-82).  For Langevin and
    Brownian systems the host draws the standard normals into
    ``arrays.random_displacement`` (randn!, src/MD.jl:4-6) unless ``noise`` is
    given."""
    ctx = _bind(arrays, parameters, neighborlist)
    _upload(arrays, neighborlist, ctx)
    xi = None
    if not isinstance(parameters.system, Newtonian):
        xi = _rng.standard_normal(arrays.r_array.shape) if noise is None else np.asarray(noise)
        arrays.random_displacement[...] = xi
    ctx.calculate_new_forces(xi)
    arrays.F_array[...] = ctx.get_state(r=False, v=False, F=True)["F"]


def do_time_step(arrays, parameters, output, neighborlist, system=None, noise=None):
    """One step, src/MD.jl:429-471 (Brownian, intended semantics) and
    :499-554 (Newtonian / Langevin)."""
    ctx = _bind(arrays, parameters, neighborlist)
    _upload(arrays, neighborlist, ctx)
    xi = None
    if not isinstance(parameters.system, Newtonian):
        xi = _rng.standard_normal(arrays.r_array.shape) if noise is None else np.asarray(noise)
        arrays.random_displacement[...] = xi.reshape(arrays.r_array.shape)
    before = ctx.rebuild_count()
    ctx.do_time_steps(1, xi)
    _sync_rebuilds(output, ctx, before)
    _download(arrays, neighborlist, ctx)


def rescale_velocities(v_array, parameters):
    """src/MD.jl:391-402 (host arithmetic on the host array, as the reference)."""
    dims = parameters.system.dims
    Ekin = 0.0
    for i in range(parameters.N):
        Ekin += float(np.sum(v_array[i] ** 2))
    Ekin *= parameters.system.m / 2.0
    kBT_current = 2 * Ekin / (dims * parameters.N)
    factor = math.sqrt(parameters.system.kBT / kBT_current)
    v_array *= factor


def calculate_full_energy(arrays, parameters, neighborlist):
    """src/AuxiliaryFunctions.jl:22-34 on the device."""
    ctx = _bind(arrays, parameters, neighborlist)
    ctx.set_state(r=arrays.r_array)
    return ctx.full_energy()


def calculate_full_energy_no_neigh(arrays, parameters, neighborlist):
    """src/AuxiliaryFunctions.jl:86-110 (O(N^2)) on the device."""
    ctx = _bind(arrays, parameters, neighborlist)
    ctx.set_state(r=arrays.r_array)
    return ctx.energy_no_neigh()


def calculate_F2_self(arrays, parameters, output, neighborlist):
    """src/AuxiliaryFunctions.jl:968-987 on the device: arrays.r_old_array is
    the reference configuration, arrays.r_array the current one (kx = 7.2).
    (The device path of perform_molecular_dynamics keeps the reference
    configuration resident and calls the context directly.)"""
    ctx = _bind(arrays, parameters, neighborlist)
    ctx.set_state(r=arrays.r_old_array)
    ctx.set_reference_positions()
    ctx.set_state(r=arrays.r_array)
    output.F2s = ctx.f2_self(7.2)


def _steinhardt(arrays, parameters, output, neighborlist, l, averaged):
    ctx = _bind(arrays, parameters, neighborlist)
    ctx.set_state(r=arrays.r_array)
    q = ctx.steinhardt(l, parameters.q_cutoff, averaged)
    if l == 6:
        output.q6 = q
    elif l == 4:
        output.q4 = q
    else:
        print(q)


def calculate_ql(arrays, parameters, output, neighborlist, l):
    """src/AuxiliaryFunctions.jl:687-711 (with calculate_qlm :758-801) on the
    device; 3D only (ScError otherwise, as the reference asserts)."""
    _steinhardt(arrays, parameters, output, neighborlist, l, False)


def calculate_ql_averaged(arrays, parameters, output, neighborlist, l):
    """src/AuxiliaryFunctions.jl:832-902 on the device."""
    _steinhardt(arrays, parameters, output, neighborlist, l, True)


def find_kinetic_energy(arrays, parameters, output):
    """src/AuxiliaryFunctions.jl:57-65."""
    v = arrays.v_array
    Ekin = float(np.sum(np.sum(v * v, axis=1)))
    output.kinetic_energy = Ekin * (parameters.system.m / 2.0)


def logdata(t):
    """src/AuxiliaryFunctions.jl:637-646."""
    if t == 0:
        return True
    log10t = math.log10(t)
    return log10t >= 3 and t % 10 ** math.floor(log10t) == 0


def create_when_to_save_array(log_factor, N_starts, N_max):
    """src/Dump.jl:18-31."""
    start_times = list(range(0, N_max + 1, max(N_max // N_starts, 1)))
    when = list(start_times)
    for i_start in start_times:
        t = 1
        while t <= N_max:
            when.append(t + i_start)
            t = math.ceil(t * log_factor)
    when.append(N_max)
    return sorted(set(x for x in when if x <= N_max))


def run_short_MD_equilibration(arrays, parameters, output, neighborlist, noise="device"):
    """src/MD.jl:579-586."""
    update_neighbor_lists(arrays, parameters, output, neighborlist)
    ctx = neighborlist.ctx
    before = ctx.rebuild_count()
    _run_device_steps(ctx, parameters, parameters.N_MD_equilibration_steps, noise, arrays)
    _sync_rebuilds(output, ctx, before)
    _download(arrays, neighborlist, ctx)


def _run_device_steps(ctx, parameters, nsteps, noise, arrays, chunk=1000):
    if nsteps <= 0:
        return
    if isinstance(parameters.system, Newtonian) or noise == "device":
        ctx.do_time_steps(nsteps, None)
        return
    done = 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        xi = _rng.standard_normal((k,) + arrays.r_array.shape)
        ctx.do_time_steps(k, xi)
        done += k


def compute_potential_energy(arrays, parameters, output, neighborlist):
    """src/AuxiliaryFunctions.jl:186-196: potential_energy_array[i] =
    find_energy(i) / 2, computed on the device."""
    ctx = _bind(arrays, parameters, neighborlist)
    arrays.potential_energy_array[...] = ctx.potential_energies()


# ------------------------------------------------------------- dump files
#
# The reference writes one HDF5 file (src/Dump.jl): groups positions/,
# velocities/, forces/, diameters/, potential_energy/, miscellaneous/ with
# one dims x N FP64 dataset per saved step, plus the parameters as file
# attributes.  No HDF5 library is available to this package, so the same
# hierarchy is kept as a directory: <filename>/<group>/<step>.npy holds the
# dataset (shape (dims, N) for vectors, element [c, i] = component c of
# particle i, as reinterpret(reshape, Float64, ...) gives it) and
# <filename>/attributes.json the attributes.

def _attr_value(x):
    if isinstance(x, (bool, int, float, str)) or x is None:
        return x
    if isinstance(x, (np.integer, np.floating, np.bool_)):
        return x.item()
    if isinstance(x, os.PathLike):
        return os.fspath(x)
    if isinstance(x, (set, frozenset, list, tuple, np.ndarray)):
        return sorted(int(v) for v in x)
    return type(x).__name__


def prepare_savefile(parameters, arrays, neighborlist=None):
    """src/Dump.jl:121-156: create the groups, store the parameter, system,
    potential and dump_info fields as attributes and the diameters.  Unlike
    the reference (whose restart then fails on the missing attribute,
    src/Restart.jl:32) when_to_save_array is stored too, and so is the skin
    distance when ``neighborlist`` is given (the reference's restart assumes
    0.35, src/Restart.jl:102)."""
    di = parameters.dump_info
    if di is None or not di.save:
        return
    root = os.fspath(di.filename)
    for g in ("positions", "velocities", "forces", "diameters", "potential_energy", "miscellaneous"):
        os.makedirs(os.path.join(root, g), exist_ok=True)
    attrs = {}
    for k, v in vars(parameters).items():
        attrs[k] = _attr_value(v) if k not in ("system", "interaction_potential", "dump_info", "callback") \
            else type(v).__name__
    for obj in (parameters.system, parameters.interaction_potential, di):
        for k, v in vars(obj).items():
            attrs[k] = _attr_value(v)
    if neighborlist is not None:
        attrs["skin_distance"] = float(neighborlist.skin_distance)
        attrs["maxneighbours"] = int(neighborlist.maxneighbours)
    with open(os.path.join(root, "attributes.json"), "w") as f:
        json.dump(attrs, f, indent=1, sort_keys=True)
    np.save(os.path.join(root, "diameters", "diameters.npy"), arrays.D_array)


def dump_writer(filename):
    """A ``writer(step, group, array)`` for save_data that stores each dataset
    as <filename>/<group>/<step>.npy."""
    def write(step, group, array):
        d = os.path.join(filename, group)
        os.makedirs(d, exist_ok=True)
        a = np.asarray(array)
        np.save(os.path.join(d, f"{int(step)}.npy"), a.T.copy() if a.ndim == 2 else a)
    return write


def save_data(arrays, parameters, output, neighborlist, restarted, writer=None):
    """src/Dump.jl:60-88: at save steps the lists are rebuilt (:66-68) and the
    requested arrays handed to ``writer(step, group, array)``; by default
    (writer None) they go to the dump directory ``dump_info.filename``."""
    di = parameters.dump_info
    if di is None or not (di.save and output.steps_done in di.when_to_save_array):
        return
    if not restarted:
        update_neighbor_lists(arrays, parameters, output, neighborlist)
    if writer is None:
        writer = dump_writer(di.filename)
    s = output.steps_done
    if di.save_r:
        writer(s, "positions", arrays.r_array.copy())
    if di.save_v:
        writer(s, "velocities", arrays.v_array.copy())
    if di.save_F:
        writer(s, "forces", arrays.F_array.copy())
    if di.save_D:
        writer(s, "diameters", arrays.D_array.copy())
    if di.save_Epot:
        compute_potential_energy(arrays, parameters, output, neighborlist)
        writer(s, "potential_energy", arrays.potential_energy_array.copy())


def read_restart(file, timestep, *, precision="fp64", cell_divisor=1):
    """src/Restart.jl:1-104: rebuild arrays, parameters, output and the
    neighbour struct from a dump at ``timestep``.  Every potential (LJ,
    Berthier) and system (Newtonian, Langevin, Brownian) is reconstructed
    (the reference handles Berthier + Newtonian only, src/Restart.jl:58-71),
    and the stored when_to_save_array and skin distance are used."""
    with open(os.path.join(file, "attributes.json")) as f:
        raw = json.load(f)

    class _Attrs(dict):  # Python stores identifiers NFKC-normalised (ϵ -> ε): look keys up the same way
        def __getitem__(self, k):
            return dict.__getitem__(self, unicodedata.normalize("NFKC", k))

        def get(self, k, d=None):
            return dict.get(self, unicodedata.normalize("NFKC", k), d)

    a = _Attrs((unicodedata.normalize("NFKC", k), v) for k, v in raw.items())
    dims, N = int(a["dims"]), int(a["N"])

    def vec(group):
        x = np.load(os.path.join(file, group, f"{int(timestep)}.npy"))
        return np.ascontiguousarray(x.T, dtype=np.float64)

    arrays = ArrayStruct(N, dims)
    arrays.r_array[...] = vec("positions")
    arrays.r_old_array[...] = arrays.r_array
    arrays.v_array[...] = vec("velocities")
    arrays.F_array[...] = vec("forces")
    arrays.D_array[...] = np.load(os.path.join(file, "diameters", "diameters.npy"))
    if a["interaction_potential"] == "Berthier":
        U = Berthier(n=int(a.get("n", 12)), rc=float(a["rc"]), σ_ratio=float(a["σ_ratio"]), ζ=float(a["ζ"]))
    elif a["interaction_potential"] == "LJ":
        U = LJ(a["ϵ"], a["σ"], a["rc"])
    else:
        raise ValueError(f"unknown interaction potential {a['interaction_potential']}")
    kind = a["system"]
    if kind == "Newtonian":
        S = Newtonian(a["kBT"], a["Δt"], a["m"], a["rescale_temperature"], dims)
    elif kind == "Langevin":
        S = Langevin(a["kBT"], a["γ"], a["Δt"], a["m"], dims)
    elif kind == "Brownian":
        S = Brownian(a["kBT"], a["γ"], a["Δt"], dims)
    else:
        raise ValueError(f"unknown system {kind}")
    di = DumpInfo(a["save"], a["filename"], set(a.get("when_to_save_array", [])), a["save_r"], a["save_v"],
                  a["save_F"], a["save_D"], a["save_Epot"])
    parameters = ParameterStruct(a["N_MD_equilibration_steps"], a["random_seed"], a["ρ"], N, a["box_size"],
                                 a["N_steps"], a["swap_probability"], a["max_MC_displacement"],
                                 a["force_cutoff2"], a["q_cutoff"], S, U, di, None)
    output = OutputStruct()
    output.steps_done = int(timestep)
    neighborlist = initialize_neighbor_struct(float(a.get("skin_distance", 0.35)), parameters.box_size,
                                              math.sqrt(parameters.force_cutoff2), N, arrays.D_array, dims,
                                              maxneighbours=int(a.get("maxneighbours", 200)),
                                              precision=precision, cell_divisor=cell_divisor)
    return arrays, parameters, output, neighborlist


ENERGY_TEST_MAX_N = 200000  # default size limit of the O(N^2) energy check at log steps


def perform_molecular_dynamics(arrays, parameters, output, neighborlist, *, restarted=False,
                               equilibrate_velocities=True, new_file=True, noise="device", writer=None,
                               log=print, energy_test=None):
    """src/MD.jl:680-719.  The step loop runs on the device in chunks between
    host-visible events: save steps (which rebuild the lists), logdata steps
    and -- if ``parameters.callback`` is not None -- every step (the callback
    then sees the downloaded state).  ``noise``: "device" (Philox on the GPU)
    or "host" (numpy normals injected, as randn! would).  ``energy_test``:
    whether log steps also compute E_pot_test, the O(N^2) all-pairs energy
    of calculate_full_energy_no_neigh (src/MD.jl:623); None (default) does so
    for N <= ENERGY_TEST_MAX_N, where one evaluation stays well under a
    second on the device.  The value is kept in output.E_pot_test."""
    arrays.r_old_array[...] = arrays.r_array
    S = parameters.system
    if not restarted and equilibrate_velocities and not isinstance(S, Brownian):
        arrays.v_array[...] = _rng.standard_normal(arrays.v_array.shape) * math.sqrt(S.kBT / S.m)
    if not restarted:
        output.steps_done = 0
    update_neighbor_lists(arrays, parameters, output, neighborlist)
    ctx = neighborlist.ctx
    ctx.set_reference_positions()  # r_old_array on the device (positions unchanged since :682)
    if equilibrate_velocities:
        run_short_MD_equilibration(arrays, parameters, output, neighborlist, noise)
        rescale_velocities(arrays.v_array, parameters)
        ctx.set_state(v=arrays.v_array)
    output.potential_energy = ctx.full_energy()
    output.N_neighbor_list_rebuilds = 0
    if new_file:  # src/MD.jl:698-701
        if writer is None:
            prepare_savefile(parameters, arrays, neighborlist)
        save_data(arrays, parameters, output, neighborlist, restarted, writer)
    if parameters.callback is not None:
        parameters.callback(arrays, parameters, output, neighborlist)
    start = time.time()
    if energy_test is None:
        energy_test = parameters.N <= ENERGY_TEST_MAX_N
    di = parameters.dump_info
    save_steps = sorted(s for s in (di.when_to_save_array if (di and di.save) else []) if s > output.steps_done)
    _upload(arrays, neighborlist, ctx)
    while output.steps_done < parameters.N_steps:
        if parameters.callback is not None:
            nxt = output.steps_done + 1
        else:
            nxt = parameters.N_steps
            for s in save_steps:
                if s > output.steps_done:
                    nxt = min(nxt, s)
                    break
            t = output.steps_done + 1
            while t < nxt and not logdata(t):
                t += 1
            nxt = min(nxt, t)
        k = nxt - output.steps_done
        before = ctx.rebuild_count()
        _run_device_steps(ctx, parameters, k, noise, arrays)
        _sync_rebuilds(output, ctx, before)
        output.steps_done += k
        _download(arrays, neighborlist, ctx)
        if parameters.callback is not None:
            parameters.callback(arrays, parameters, output, neighborlist)
            _upload(arrays, neighborlist, ctx)
        save_data(arrays, parameters, output, neighborlist, restarted, writer)
        if logdata(output.steps_done):
            # print_log_data (src/MD.jl:610-650) with its observables on the device
            output.potential_energy = ctx.full_energy()
            if not isinstance(S, Brownian):
                find_kinetic_energy(arrays, parameters, output)
            else:
                output.kinetic_energy = 0.0
            if parameters.q_cutoff > 0.0:
                output.q6 = ctx.steinhardt(6, parameters.q_cutoff, True)
                output.q4 = ctx.steinhardt(4, parameters.q_cutoff, True)
            output.F2s = ctx.f2_self(7.2)
            output.E_pot_test = ctx.energy_no_neigh() if energy_test else None
            if log is not None:
                kT = output.kinetic_energy * 2 / S.dims / parameters.N
                qs = f"q4 = {output.q4:.3f}, q6 = {output.q6:.3f}, " if S.dims == 3 else ""
                et = f"E_pot_test = {output.E_pot_test:.12f}, " if energy_test else ""
                log(f"{output.steps_done}/{parameters.N_steps}, "
                    f"E = {output.potential_energy + output.kinetic_energy:.2f}, "
                    f"E_pot = {output.potential_energy:.12f}, {et}E_kin = {output.kinetic_energy:.2f}, "
                    f"kT = {kT:.3f}, {qs}F2s = {output.F2s:.3f}, elapsed = {time.time() - start:.2f}s")
            if output.q6 > 0.2:  # src/MD.jl:711-713
                raise RuntimeError("Simulation has crystallized, terminating...")
    if log is not None:
        log(f"Number of neighbor lists built = {output.N_neighbor_list_rebuilds}")


def find_inherent_structure(r_array_to_be_minimized, parameters, arrays, neighborlist, output, *,
                            tolerance=1e-6, max_iterations=10 ** 6, max_step_size=0.001, verbose=1, dt=0.005):
    """src/InherentStructure.jl:4-87: capped steepest descent on the device
    (pair forces only) until max(max|F|, step) <= tolerance, checked after
    every iteration's move as the reference checks it; RuntimeError("Not
    converged") past max_iterations.  arrays.r_array, F_array (the last
    iteration's forces) and potential_energy_array are mutated, skin-triggered
    rebuilds are counted in output.  verbose >= 1 prints the reference's
    closing line (iterations, total potential energy, largest displacement);
    verbose >= 2 also its report after iterations 1, 1001, 2001, ...
    (src/InherentStructure.jl:57-69): the descent then runs as device calls
    of 1 and then 1000 iterations, which take the same iterations as one
    call (the stop test is made after every iteration either way)."""
    r0 = np.array(r_array_to_be_minimized, dtype=np.float64, copy=True)
    arrays.r_array[...] = r0
    update_neighbor_lists(arrays, parameters, output, neighborlist)
    ctx = neighborlist.ctx
    before = ctx.rebuild_count()
    L = parameters.box_size

    def maxchange():
        d = arrays.r_array - r0
        d -= np.round(d / L) * L
        return float(np.sqrt(np.max(np.sum(d * d, axis=1)))) if len(d) else 0.0

    if verbose >= 2:
        it, converged, chunk = 0, False, 1
        while it < max_iterations and not converged:
            k = min(chunk, max_iterations - it)
            done, mf, converged = ctx.minimize_until(k, dt, max_step_size, tolerance)
            it += done
            if done == k and it % 1000 == 1:
                arrays.r_array[...] = ctx.get_state(r=True, v=False, F=False)["r"]
                compute_potential_energy(arrays, parameters, output, neighborlist)
                print("(iteration, maxchange, stepsize, maxforce, sum(arrays.potential_energy_array)) = ",
                      (it, maxchange(), mf * dt, mf, float(np.sum(arrays.potential_energy_array))), sep="")
            chunk = 1000
    else:
        it, _, converged = ctx.minimize_until(max_iterations, dt, max_step_size, tolerance)
    _sync_rebuilds(output, ctx, before)
    _download(arrays, neighborlist, ctx)
    if not converged:
        raise RuntimeError("Not converged")
    compute_potential_energy(arrays, parameters, output, neighborlist)
    if verbose >= 1:
        print("after ", it, " iterations, the potential energy is ", float(np.sum(arrays.potential_energy_array)),
              ". Maximal total displacement = ", maxchange(), sep="")
    return arrays.r_array.copy()
