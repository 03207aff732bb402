"""Benchmark: particle-steps/s of the MD hot path on one MI355X.

Default workload (the large single system): 3D polydisperse non-additive soft
spheres, N = 10^6, rho = 1.02, skin 0.3, Newtonian NVE (velocity Verlet,
dt = 0.005) from a seeded jittered simple-cubic lattice, relaxed on the
device by capped steepest descent and thermalised at kBT = 0.2 by velocity
rescaling (all untimed).  Inputs are resident in HBM when the timed region
starts.  The timed region is K steps of do_time_step (neighbour rebuilds on
the skin trigger included) in one sc_do_time_steps call, bracketed by device
synchronisation.  `value` is the FP64 drop-in form; the mixed-precision form
(FP32 pair arithmetic) runs from the same state and is reported beside it
with its force error and NVE drift (linear fit of the total energy over a
separate run, scaled to 10^5 steps).

roofline: the dominant kernel is the pair-force kernel (fused with the
second half-kick and the COM partial sums).  Algorithmic bytes per particle:
r (24) + D (8) + v read/write (48) + F write (24) + 4 (n_half+1) neighbour
index bytes, n_half measured on the device lists; its average duration comes
from HIP events on the context stream over the timed region.

cpu_baseline: the oracle (FP64 C restatement of the reference's serial
algorithm, one thread) on the same N = 10^6 state, timed for a bounded number
of steps (~15 s), rebuilds included.

--workload batched: 256 independent replicas (N log-uniform in [500, 20000],
75 % 3D rho 1.02 / 25 % 2D rho 1.0, kBT in [0.1, 0.3], Brownian, per-replica
seeds, mixed precision) split over two batches stepped concurrently on their
own HIP streams, kernels launched eagerly (a captured hipGraph replay and a
single stream were both measured slower); value = aggregate
particle-steps/s over the wall time of the timed steps.  dt = 1e-4: at the stated dt = 0.001 explicit
Euler-Maruyama blows up within 10^2-10^3 steps for this potential.
--workload served: see simulationcode_amd/serve.py.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic(name):
    """HBM bytes per launch of a kernel from a committed counter summary
    (profiles/<name>.txt, line 'bytes_per_launch = B': FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by a copy kernel's reading), or None
    when no such summary exists."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch"):
            return float(line.split("=")[1])
    return None
METRIC = "particle-steps/s FP64 vs mixed @ stated force err; p50/p99 replica latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_ctx(nat, N, L, D, precision, dt=0.005, kBT=0.2, seed=1, cell_divisor=0):
    from simulationcode_amd import md
    U = md.Berthier(rc=1.25)
    p = nat.ScParams()
    p.dims, p.potential, p.system, p.precision = 3, nat.POT_BERTHIER, nat.SYS_NEWTONIAN, precision
    p.rescale_temperature, p.cell_divisor = 0, cell_divisor
    p.N, p.maxneighbours = N, 200
    p.box_size, p.force_cutoff, p.force_cutoff2, p.skin_distance = L, 1.25, 1.25 ** 2, 0.3
    p.b_c0, p.b_c2, p.b_c4, p.b_zeta = U.c0, U.c2, U.c4, U.ζ
    p.kBT, p.gamma, p.dt, p.m, p.seed = kBT, 1.0, dt, 1.0, seed
    return nat.Context(p, D), p


def total_energy(ctx):
    return ctx.full_energy() + ctx.kinetic_energy()


def timed_run(ctx, steps, warmup):
    ctx.do_time_steps(warmup)
    reb0 = ctx.rebuild_count()
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    ctx.do_time_steps(steps)  # synchronises on return
    wall = time.perf_counter() - t0
    kt = ctx.kernel_times()
    ctx.set_kernel_timing(False)
    return {"wall_s": wall, "device_ms": ctx.last_elapsed_ms(), "rebuilds": ctx.rebuild_count() - reb0, "kt": kt}


def drift_run(ctx, steps, samples=40):
    """Total-energy drift: least-squares slope of E over `steps` steps,
    relative to |E|, scaled to 10^5 steps; also the relative spread."""
    chunk = max(steps // samples, 1)
    t, e = [0], [total_energy(ctx)]
    for k in range(samples):
        ctx.do_time_steps(chunk)
        t.append((k + 1) * chunk)
        e.append(total_energy(ctx))
    slope = np.polyfit(np.array(t, float), np.array(e), 1)[0]
    return float(slope * 1e5 / abs(e[0])), float((max(e) - min(e)) / abs(e[0]))


def cpu_baseline(op_state, budget_s=15.0):
    import oracle
    r, v, F, D, L, N = op_state
    op = oracle.make_params(dims=3, N=N, box_size=L, force_cutoff=1.25, skin=0.3, D=D,
                            potential=oracle.POT_BERTHIER, system=oracle.SYS_NEWTONIAN, dt=0.005, kBT=0.2,
                            maxneighbours=128)
    st = oracle.State(op, r, v, D, F=F)
    st.update_neighbor_lists(cells=True)
    steps = 0
    t0 = time.perf_counter()
    while True:
        st.do_time_steps(1, cells=True)
        steps += 1
        el = time.perf_counter() - t0
        if (el > budget_s and steps >= 3) or steps >= 2000:
            break
    out = {"value": N * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
           "sample": f"oracle (serial FP64 C restatement, cell-list rebuild) on the same N={N} state, "
                     f"{steps} steps in {el:.1f} s incl. {int(st.rebuilds[0]) - 1} rebuilds"}
    # the same step on all host cores (OpenMP: per-particle rows, gathered
    # forces); a 3-step pilot sizes one run of about budget_s
    threads = min(16, len(os.sched_getaffinity(0)))
    rb, vb, Fb = (np.ascontiguousarray(x, dtype=np.float64).copy() for x in (r, v, F))
    t0 = time.perf_counter()
    oracle.baseline_steps(op, 3, threads, rb, vb, Fb, D)
    per = (time.perf_counter() - t0) / 3
    n = int(min(max(budget_s / max(per, 1e-6), 5), 5000))
    t0 = time.perf_counter()
    nreb = oracle.baseline_steps(op, n, threads, rb, vb, Fb, D)
    el = time.perf_counter() - t0
    out["all_cores"] = {"value": N * n / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                        "sample": f"the same Newtonian step on {threads} OpenMP threads (per-particle rows, "
                                  f"gathered forces), {n} steps in {el:.1f} s incl. {nreb + 1} rebuilds"}
    return out


def bench_large(args, nat, synth):
    N = args.N
    rho, dims = 1.02, 3
    L = synth.box_size(N, rho, dims)
    rng = np.random.default_rng(args.seed)
    D = synth.berthier_diameters(N, rng)
    r = synth.lattice_positions(N, dims, L, 0.1, rng)

    log(f"[bench] setup N={N} L={L:.3f}")
    cc = None if args.no_copy_ceiling else nat.copy_bandwidth()  # streaming ceiling, 1 GiB copies
    ctx, p = make_ctx(nat, N, L, D, nat.PREC_FP64, cell_divisor=args.cell_divisor)
    ctx.set_state(r=r, v=np.zeros((N, dims)), F=np.zeros((N, dims)))
    ctx.update_neighbor_lists()
    t0 = time.perf_counter()
    fmax = ctx.minimize(400, 0.005, 0.02)
    log(f"[bench] relaxed: max|F| = {fmax:.3g} ({time.perf_counter() - t0:.1f}s)")
    v = synth.maxwell_boltzmann(N, dims, 0.2, 1.0, rng)
    ctx.set_state(v=v, F=np.zeros((N, dims)))
    ctx.update_neighbor_lists()
    ctx.calculate_new_forces(None)
    for _ in range(10):
        ctx.do_time_steps(50)
        ctx.rescale_velocities()
    log("[bench] thermalised at kBT=0.2")
    state = ctx.get_state()
    tot, mx = ctx.neighbor_stats()
    n_half = tot / 2.0 / N

    res64 = timed_run(ctx, args.steps, args.warmup)
    ms64 = res64["device_ms"] / args.steps
    log(f"[bench] fp64: {ms64 * 1e3:.1f} us/step, rebuilds {res64['rebuilds']}")
    value = N * args.steps / (res64["device_ms"] / 1e3)

    kt = res64["kt"]
    f_avg_s = kt["force_ms"] / max(kt["force_launches"], 1) / 1e3
    bytes_per_particle = 24 + 8 + 48 + 24 + 4 * (n_half + 1)
    achieved = N * bytes_per_particle / f_avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("traffic_k_force1_berthier3d_N1e6_fp64"),
                "kernel": "k_force<1> (pair forces + second half-kick + COM partials)",
                "kernel_avg_us": round(f_avg_s * 1e6, 2), "bytes_per_particle": round(bytes_per_particle, 2),
                "n_half_mean": round(n_half, 3),
                "drift_kernel_avg_us": round(kt["drift_ms"] / max(kt["drift_launches"], 1) * 1e3, 2)}
    if cc:
        roofline.update({"copy_ceiling_GBs": round(cc, 1), "frac_of_copy_ceiling": round(achieved / cc, 4)})
        log(f"[bench] measured copy ceiling {cc:.0f} GB/s")
    out = {"metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded jittered lattice, relaxed and thermalised on device)",
           "config": {"workload": f"3D polydisperse non-additive soft spheres, N={N}, rho=1.02, Newtonian NVE "
                                  f"dt=0.005, skin 0.3, kBT=0.2 start", "N": N, "dims": 3,
                      "potential": "Berthier (rc=1.25, zeta=0.2)", "dynamics": "Newtonian", "dt": 0.005},
           "roofline": roofline, "fp64": {"rebuilds": res64["rebuilds"], "wall_s": res64["wall_s"]}}

    if args.drift_steps > 0:
        ctx.set_state(r=state["r"], v=state["v"], F=state["F"])
        ctx.update_neighbor_lists()
        d64, fl64 = drift_run(ctx, args.drift_steps)
        out["fp64"].update({"nve_rel_drift_per_1e5_steps": d64, "nve_rel_fluct": fl64,
                            "drift_window_steps": args.drift_steps})

    if not args.no_mixed:
        ctxm, _ = make_ctx(nat, N, L, D, nat.PREC_MIXED, cell_divisor=args.cell_divisor)
        ctxm.set_state(r=state["r"], v=state["v"], F=state["F"])
        ctxm.update_neighbor_lists()
        ctx.set_state(r=state["r"], v=state["v"], F=state["F"])
        ctx.update_neighbor_lists()
        ctx.calculate_new_forces(None)
        ctxm.calculate_new_forces(None)
        F64 = ctx.get_state(F=True, r=False, v=False)["F"]
        F32 = ctxm.get_state(F=True, r=False, v=False)["F"]
        fn = np.linalg.norm(F64, axis=1)
        dn = np.linalg.norm(F32 - F64, axis=1)
        ctx.close()
        ctxm.set_state(r=state["r"], v=state["v"], F=state["F"])
        ctxm.update_neighbor_lists()
        resm = timed_run(ctxm, args.steps, args.warmup)
        msm = resm["device_ms"] / args.steps
        kmt = resm["kt"]
        out["mixed"] = {"value": N * args.steps / (resm["device_ms"] / 1e3), "ms_per_step": msm,
                        "speedup_vs_fp64": ms64 / msm,
                        "max_rel_force_err_vs_mean": float(dn.max() / fn.mean()),
                        "max_rel_force_err_vs_max": float(dn.max() / fn.max()),
                        "rebuilds": resm["rebuilds"],
                        "force_kernel_avg_us": kmt["force_ms"] / max(kmt["force_launches"], 1) * 1e3}
        if args.drift_steps > 0:
            ctxm.set_state(r=state["r"], v=state["v"], F=state["F"])
            ctxm.update_neighbor_lists()
            dm, flm = drift_run(ctxm, args.drift_steps)
            out["mixed"].update({"nve_rel_drift_per_1e5_steps": dm, "nve_rel_fluct": flm})
        log(f"[bench] mixed: {msm * 1e3:.1f} us/step")
        ctxm.close()
    else:
        ctx.close()

    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, 1 thread and all host cores)")
        out["cpu_baseline"] = cpu_baseline((state["r"], state["v"], state["F"], D, L, N))
    else:
        out["cpu_baseline"] = None
    return out


def bench_batched(args, nat):
    from simulationcode_amd import workloads as W
    specs = W.batched_specs(args.replicas)
    t0 = time.perf_counter()
    if args.cell_divisor:
        shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else (256, 4)
        cd, tune = args.cell_divisor, None
    else:
        cd, shape, tune = W.autotune_batch(specs, pilot=64, steps=60)
    log(f"[bench] autotuner: cell divisor {cd}, build block {shape[0]} x {shape[1]} lanes/slot "
        f"({time.perf_counter() - t0:.1f}s)")
    if args.streams > 1:
        b = nat.BatchGroup(len(specs) + 1, streams=args.streams, graph=args.graph)
    else:
        b = nat.Batch(len(specs) + 1, graph=args.graph)
    b.set_build_launch(*shape)
    ids, Ns, dims = [], [], []
    t0 = time.perf_counter()
    for s in specs:
        p, D, r, v = W.replica(**dict(s, cell_divisor=cd))
        ids.append(b.add(p, D, r, v))
        Ns.append(s["N"])
        dims.append(s["dims"])
    log(f"[bench] admitted {len(ids)} replicas, {sum(Ns)} particles ({time.perf_counter() - t0:.1f}s)")
    b.minimize(300, 0.005, 0.02)
    b.step(args.warmup)
    full_tune = None
    if not args.cell_divisor:
        # the launch shape changes no result: re-tune it on the full batch
        # (the pilot holds a quarter of the replicas, so fewer rebuild per step)
        tab, best = b.autotune(100)
        shape = (tab[best]["block_size"], tab[best]["lanes_per_particle"])
        full_tune = [dict(e, ms_per_step=round(e["ms_per_step"], 5)) for e in tab]
        log(f"[bench] full-batch shape {shape[0]} x {shape[1]} lanes/slot")
    b.step(args.steps)
    ms = b.info()["last_ms"]
    tot = sum(Ns)
    value = tot * args.steps / (ms / 1e3)
    # algorithmic bytes per particle-step, Brownian with on-device noise:
    # r 2*8d + F write 8d + disp 2*8d + D 8 + 4 (n_half + 1)
    byt = 0.0
    for rid, n, d in zip(ids, Ns, dims):
        t, _ = b.neighbor_stats(rid)
        byt += n * (40 * d + 8 + 4 * (t / 2.0 / n + 1))
    achieved = byt * args.steps / (ms / 1e3) / 1e9
    log(f"[bench] batched: {ms / args.steps * 1e3:.1f} us/step over {tot} particles")
    return {"metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "mixed (f32 pair / f64 state)",
            "data": "synthetic (seeded lattices relaxed on device)",
            "config": {"workload": f"{len(ids)} independent polydisperse replicas, N log-uniform [500, 20000], "
                                   "75% 3D / 25% 2D, Brownian dt=1e-4, kBT U[0.1,0.3], per-replica seeds",
                       "replicas": len(ids), "particles": tot, "graph": args.graph, "streams": args.streams,
                       "cell_divisor": cd, "build_block": shape[0], "build_lanes_per_slot": shape[1],
                       "autotune": tune, "autotune_full_batch": full_tune},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "kernel": "whole batched step (all kernels)",
                         "bytes_per_particle_step_mean": round(byt / tot, 2)},
            "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--N", type=int, default=10 ** 6)
    ap.add_argument("--workload", choices=("large", "batched", "served", "copycal"), default="large",
                    help="copycal: five plain 256 MiB device copies, the reference reading for HBM counters")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-copy-ceiling", action="store_true", help="large: skip the measured device-copy ceiling")
    ap.add_argument("--graph", action="store_true",
                    help="batched: replay the step as a captured hipGraph (eager launches measured faster)")
    ap.add_argument("--streams", type=int, default=2,
                    help="batched: split the replicas over this many batches on their own streams")
    ap.add_argument("--drift-steps", type=int, default=20000)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261206)
    ap.add_argument("--cell-divisor", type=int, default=0,
                    help="0: library default (large) / autotuned (batched)")
    ap.add_argument("--shape", default="", help="batched with --cell-divisor: build launch shape BS,LPP")
    ap.add_argument("--rate", type=float, nargs="*", default=None, help="served: arrival rates (jobs/s)")
    ap.add_argument("--jobs", type=int, default=48, help="served: jobs per arrival rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the MD hot path here runs on a single GPU "
              "(replicas would shard across GPUs as independent batches, a single large system by "
              "spatial domain decomposition; neither is built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copycal":
        import torch
        a = torch.empty(1 << 26, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copycal: 5 copies of 256 MiB")
        return
    from simulationcode_amd import _native as nat, synth
    nat.load()
    if args.workload == "large":
        out = bench_large(args, nat, synth)
    elif args.workload == "batched":
        out = bench_batched(args, nat)
    else:
        from simulationcode_amd import serve
        out = serve.bench(args, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
