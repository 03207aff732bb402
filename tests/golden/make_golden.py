"""Writes the golden fixtures under tests/golden/ with the oracle.

The reference holds no golden vectors (it has no tests); these fixtures pin
the oracle against regressions and give the GPU tests inputs/outputs that do
not depend on the oracle build at run time.  Each fixture: seeded inputs
(r, v, D, injected noise, parameters) and the oracle's outputs (half/full
pair sets, forces, energies, state after 50 steps, rebuild count).

The minimiser fixtures hold an overlapping start and the oracle's
find_inherent_structure result (iteration count, positions, last forces,
rebuild count).

Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests import helpers as H  # noqa: E402

FIXTURES = [
    # name, kind, dims, N, system
    ("lj3d_newtonian_N216", "lj", 3, 216, oracle.SYS_NEWTONIAN),
    ("berthier3d_langevin_N343", "berthier", 3, 343, oracle.SYS_LANGEVIN),
    ("berthier2d_brownian_N256", "berthier", 2, 256, oracle.SYS_BROWNIAN),
    ("lj2d_langevin_N196", "lj", 2, 196, oracle.SYS_LANGEVIN),
]
STEPS = 50
PARAM_FIELDS = [f for f, _ in oracle.OrParams._fields_]
# find_inherent_structure from an overlapping (unrelaxed) start:
# name, dims, N, tolerance, max_iterations (dt 0.005, max step 0.001)
MINIMISER_FIXTURES = [
    ("minimiser_berthier3d_N125", 3, 125, 1e-4, 100000),
    ("minimiser_berthier2d_N144", 2, 144, 1e-4, 100000),
]
MIN_DT, MIN_MAX_STEP = 0.005, 0.001


def make(name, kind, dims, N, system):
    dt = 0.001 if system == oracle.SYS_BROWNIAN else 0.005
    op, r, D, v, rc = H.make_system(kind, N, dims, seed=sum(name.encode()) * 7 + N, system=system, dt=dt,
                                    kBT=0.2)
    rng = np.random.default_rng(N * 31 + dims)
    noise = rng.standard_normal((STEPS + 1, N, dims)) if system != oracle.SYS_NEWTONIAN else None
    lists = oracle.build_lists(op, r, D, cells=False)
    F0 = oracle.calculate_new_forces(op, r, D, lists, v, None if noise is None else noise[0])
    st = oracle.State(op, r, v, D, F=F0)
    st.update_neighbor_lists(cells=False)
    st.do_time_steps(STEPS, None if noise is None else noise[1:], cells=False)
    params = {f: getattr(op, f) for f in PARAM_FIELDS}
    params["force_cutoff"] = rc
    out = dict(r=r, v=v, D=D, full_pairs=lists.full_pairs(), half_pairs=lists.half_pairs(), nfull=lists.nfull,
               F0=F0, energy=oracle.full_energy(op, r, D, lists), energy_nn=oracle.energy_no_neigh(op, r, D),
               ekin=oracle.kinetic_energy(op, v), r_end=st.r, v_end=st.v, F_end=st.F, disp_end=st.disp,
               rebuilds=st.rebuilds[0], steps=STEPS, params=json.dumps(params))
    if noise is not None:
        out["noise"] = noise
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def make_minimiser(name, dims, N, tol, max_it):
    op, r, D, v, rc = H.make_system("berthier", N, dims, seed=31, relax=False)
    st = oracle.State(op, r, v, D)
    it, fm, ok = st.find_inherent_structure(tolerance=tol, max_iterations=max_it, max_step=MIN_MAX_STEP, dt=MIN_DT)
    assert ok
    params = {f: getattr(op, f) for f in PARAM_FIELDS}
    params["force_cutoff"] = rc
    np.savez_compressed(os.path.join(HERE, name + ".npz"), r=r, v=v, D=D, tolerance=tol, max_iterations=max_it,
                        dt=MIN_DT, max_step=MIN_MAX_STEP, iterations=it, maxforce=fm, r_min=st.r, F_last=st.F,
                        rebuilds=st.rebuilds[0], params=json.dumps(params))


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["params"] = json.loads(str(d["params"]))
    return d


def params_of(d):
    p = oracle.OrParams()
    for f in PARAM_FIELDS:
        setattr(p, f, d["params"][f])
    return p


if __name__ == "__main__":
    for fx in FIXTURES:
        make(*fx)
        print("wrote", fx[0])
    for fx in MINIMISER_FIXTURES:
        make_minimiser(*fx)
        print("wrote", fx[0])
