"""Replica specifications for the batched and served workloads.

A replica is one independent periodic system handed to libsc: sc_params,
diameters, positions and velocities, generated from a seed (lattice + jitter
positions, the polydisperse diameter recipe of src/InitialConfiguration.jl:1-11,
Maxwell-Boltzmann velocities).

Batched workload: 256 replicas, N log-uniform in [500, 20000], 3D with
probability 0.75 (rho = 1.02) else 2D (rho = 1.0), kBT uniform in [0.1, 0.3],
Brownian dt = 0.001, gamma = 1, per-replica seeds, mixed precision.
Served workload: Poisson arrivals of jobs with N log-uniform in [10^3, 2 10^5],
dims 2 or 3, dynamics uniform over Newtonian / Langevin / Brownian, steps
log-uniform in [10^3, 10^5], polydisperse potential.
"""
import math

import numpy as np

from . import _native as nat
from . import synth
from .md import Berthier, LJ

SYSTEMS = {"newtonian": nat.SYS_NEWTONIAN, "langevin": nat.SYS_LANGEVIN, "brownian": nat.SYS_BROWNIAN}


def replica(N, dims, *, potential="berthier", system="brownian", kBT=0.2, dt=0.001, gamma=1.0, m=1.0,
            seed=0, precision="mixed", rho=None, skin=0.3, cell_divisor=0, build_mode=0, maxneighbours=200,
            jitter=0.1):
    """(ScParams, D, r, v) of one seeded replica."""
    rng = np.random.default_rng(seed)
    p = nat.ScParams()
    p.dims = dims
    p.system = SYSTEMS[system] if isinstance(system, str) else int(system)
    p.precision = nat.PREC_MIXED if precision == "mixed" else nat.PREC_FP64
    p.cell_divisor = cell_divisor
    p.build_mode = build_mode
    p.N = N
    p.maxneighbours = maxneighbours
    if potential == "berthier":
        rho = rho or (1.02 if dims == 3 else 1.0)
        U = Berthier(rc=1.25)
        p.potential = nat.POT_BERTHIER
        p.b_c0, p.b_c2, p.b_c4, p.b_zeta = U.c0, U.c2, U.c4, U.ζ
        rc = 1.25
        D = synth.berthier_diameters(N, rng)
    else:
        rho = rho or 1.0
        U = LJ(1.0, 1.0, 2.5)
        p.potential = nat.POT_LJ
        p.lj_eps, p.lj_sigma, p.lj_Uc = U.ϵ, U.σ, U.Uc
        rc = 2.5
        D = np.ones(N)
    L = synth.box_size(N, rho, dims)
    p.box_size = L
    p.force_cutoff = rc
    p.force_cutoff2 = rc * rc
    p.skin_distance = skin
    p.kBT, p.gamma, p.dt, p.m = kBT, gamma, dt, m
    p.seed = (int(seed) * 0x9E3779B97F4A7C15 + 1) & 0xFFFFFFFFFFFFFFFF
    r = synth.lattice_positions(N, dims, L, jitter, rng)
    if p.system == nat.SYS_BROWNIAN:
        v = np.zeros((N, dims))
    else:
        v = synth.maxwell_boltzmann(N, dims, kBT, m, rng)
    return p, D, r, v


def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def batched_specs(n_replicas=256, seed=20261207, n_lo=500, n_hi=20000, dt=1e-4):
    """Keyword dicts for replica() of the batched workload.  The workload is
    stated with dt = 0.001, at which explicit Euler-Maruyama blows up within
    10^2-10^3 steps for the r^-12 core (observed in FP64 and mixed alike);
    dt = 1e-4 is used instead (pass dt=0.001 to reproduce)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_replicas):
        dims = 3 if rng.uniform() < 0.75 else 2
        out.append(dict(N=log_uniform_int(rng, n_lo, n_hi), dims=dims, system="brownian",
                        kBT=float(rng.uniform(0.1, 0.3)), dt=dt, gamma=1.0, seed=seed * 1000 + k,
                        precision="mixed"))
    return out


def served_job(rng, job_id, n_lo=1000, n_hi=200000, steps_lo=1000, steps_hi=100000):
    """One served job: replica() keywords plus its step count."""
    dims = 2 if rng.uniform() < 0.5 else 3
    system = ("newtonian", "langevin", "brownian")[int(rng.integers(0, 3))]
    # explicit Euler-Maruyama is only conditionally stable for the r^-12 core:
    # at dt = 0.001 rare deep overlaps blow up within ~10^3 steps, so the long
    # served Brownian jobs use dt = 1e-4 (velocity Verlet is stable at 0.005)
    dt = 1e-4 if system == "brownian" else 0.005
    return dict(N=log_uniform_int(rng, n_lo, n_hi), dims=dims, system=system, kBT=float(rng.uniform(0.1, 0.3)),
                dt=dt, gamma=1.0, seed=int(rng.integers(1, 2 ** 62)), precision="mixed"), \
        log_uniform_int(rng, steps_lo, steps_hi)


# build launch shapes tried by autotune_group: one lane per slot (the radius
# form, the R_v sweep) and the R_v sweep with several lanes per slot
GROUP_SHAPES = ((256, 0), (256, 1), (256, 2), (256, 4), (256, 8), (256, 16), (64, 4), (64, 8))


def autotune_group(specs, cell_divisors=(1, 2, 3), shapes=GROUP_SHAPES, streams=4, chunk=12, steps=120,
                   relax=(300, 0.005, 0.02), warmup=20):
    """Workload-level autotuner that measures every choice as it runs: for
    each cell divisor, all replicas of `specs` are admitted into `streams`
    concurrently stepped batches (steps launched ahead in chunks of `chunk`),
    relaxed and warmed up; then each build launch shape steps the whole group
    `steps` steps and its wall time per step is recorded.  The state advances
    through the measurements (every shape writes the same lists, so the
    trajectory does not depend on the order).  Returns (best cell divisor,
    best shape, table).  The cell divisor sets the slot order and so the bits
    of a trajectory: it is applied to every replica of a run alike."""
    table = []
    for cd in cell_divisors:
        b = nat.BatchGroup(len(specs) + 1, streams=streams, graph=False) if streams > 1 \
            else nat.Batch(len(specs) + 1, graph=False)
        b.set_ahead(chunk)
        for spec in specs:
            p, D, r, v = replica(**dict(spec, cell_divisor=cd))
            b.add(p, D, r, v)
        b.minimize(*relax)
        b.step(warmup)
        for shape in shapes:
            b.set_build_launch(*shape)
            b.step(steps)
            ms = b.info()["last_ms"]
            table.append({"cell_divisor": cd, "block_size": shape[0], "lanes_per_particle": shape[1],
                          "ms_per_step": ms / steps})
        b.close()
    best = min(table, key=lambda e: e["ms_per_step"])
    return best["cell_divisor"], (best["block_size"], best["lanes_per_particle"]), table

