"""simulationcode_amd -- MI355X-native molecular-dynamics hot path of
SimulationCode.jl (cell-list Verlet rebuild, LJ / Berthier pair forces and
energies, Newtonian / Langevin / Brownian integrators) as hand-written HIP
kernels for gfx950 behind the C-ABI in include/sc.h.

``simulationcode_amd.md`` mirrors the reference's Julia API; the
``_native`` module is the ctypes binding of libsc.
"""
from . import _native, synth  # noqa: F401
from .md import (  # noqa: F401
    LJ, Berthier, Brownian, Newtonian, Langevin, DumpInfo, ParameterStruct, ArrayStruct, OutputStruct,
    NeighborListStruct, initialize_neighbor_struct, update_neighbor_lists, calculate_new_forces,
    do_time_step, rescale_velocities, calculate_full_energy, calculate_full_energy_no_neigh,
    find_kinetic_energy, logdata, create_when_to_save_array, run_short_MD_equilibration,
    perform_molecular_dynamics, find_inherent_structure, save_data, seed, compute_potential_energy,
    prepare_savefile, dump_writer, read_restart,
)
