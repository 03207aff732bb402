"""Benchmark: particle-steps/s of the MD hot path on one MI355X.

Default workload (the large single system): 3D polydisperse non-additive soft
spheres, N = 10^6, rho = 1.02, skin 0.3, Newtonian NVE (velocity Verlet,
dt = 0.005) from a seeded jittered simple-cubic lattice, relaxed on the
device by capped steepest descent and thermalised at kBT = 0.2 by velocity
rescaling (all untimed).  Inputs are resident in HBM when the timed region
starts.

Each precision form (FP64, mixed) starts from the same thermalised state:
update_neighbor_lists!, W untimed warm-up steps, then the K timed steps in one
sc_do_time_steps call with per-kernel timing off (device time from HIP
events around the call; the call synchronises).  `value` is the FP64 form's
rate over those K steps; each form also reports its rebuild count in the K
steps, a separate run with per-kernel timing on (force and drift launch
averages), its steady step without rebuilds, and a long window of
--long-steps steps (>= 5 rebuild intervals) whose ratio is the like-for-like
mixed / FP64 speed-up; the mixed form's force error on the common state and
both forms' NVE drift (linear fit of the total energy over a separate run,
scaled to 10^5 steps).

roofline: the dominant kernel is the pair-force kernel (fused with the
second half-kick and the COM partial sums).  Algorithmic bytes per particle:
r (24) + D (8) + v read/write (48) + F write (24) + 4 (n_half+1) neighbour
index bytes, n_half measured on the device lists; its average launch time
comes from HIP events on the context stream in the timed-kernel run.

small_configs: the two small stated systems (LJ N=1000 Newtonian; polydisperse
N=8000 Langevin, FP64 and mixed), see small_configs.

batched / served: bounded runs of the replica workloads (extra keys), see
batched_summary and simulationcode_amd/serve.summary.

cpu_baseline: the oracle (FP64 C restatement of the reference's serial
algorithm) on the same N = 10^6 state on the host: 10 warm-up steps, then the
median of three runs, once on one thread and once on all host cores.

--workload batched: 256 independent replicas (N log-uniform in [500, 20000],
75 % 3D rho 1.02 / 25 % 2D rho 1.0, kBT in [0.1, 0.3], Brownian, per-replica
seeds, mixed precision) split over three batches stepped concurrently on their
own HIP streams, steps launched ahead in chunks of 12 (a captured hipGraph
replay, one, two or five to eight streams and chunks of 4-16 steps were all
measured slower; four streams were 2% faster only when no other stream,
the default stream included, had been created in the process before them:
with four hardware queues per process, five streams share one); value = aggregate
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


def make_ctx(nat, N, L, D, precision, dt=0.005, kBT=0.2, seed=1, cell_divisor=0, build_mode=0):
    from simulationcode_amd import md
    U = md.Berthier(rc=1.25)
    p = nat.ScParams()
    p.dims, p.potential, p.system, p.precision = 3, nat.POT_BERTHIER, nat.SYS_NEWTONIAN, precision
    p.rescale_temperature, p.cell_divisor, p.build_mode = 0, cell_divisor, build_mode
    p.N, p.maxneighbours = N, 200
    p.box_size, p.force_cutoff, p.force_cutoff2, p.skin_distance = L, 1.25, 1.25 ** 2, 0.3
    p.b_c0, p.b_c2, p.b_c4, p.b_zeta = U.c0, U.c2, U.c4, U.ζ
    p.kBT, p.gamma, p.dt, p.m, p.seed = kBT, 1.0, dt, 1.0, seed
    return nat.Context(p, D), p


def total_energy(ctx):
    return ctx.full_energy() + ctx.kinetic_energy()


def drift_run(ctx, steps, samples=40):
    """Total-energy drift: least-squares slope of E over `steps` steps,
    relative to |E|, scaled to 10^5 steps; also the relative spread."""
    chunk = max(steps // samples, 1)
    t, e = [0], [total_energy(ctx)]
    for k in range(samples):
        ctx.do_time_steps(chunk)
        t.append((k + 1) * chunk)
        e.append(total_energy(ctx))
        if steps >= 200000:  # a long window reports as it goes
            log(f"[bench] drift window: {t[-1]} of {steps} steps, E = {e[-1]:.10g}")
    slope = np.polyfit(np.array(t, float), np.array(e), 1)[0]
    return float(slope * 1e5 / abs(e[0])), float((max(e) - min(e)) / abs(e[0]))


def leg(ctx, state, args):
    """One precision form from the common state: the forced rebuild of
    update_neighbor_lists!, W untimed warm-up steps, then the K headline steps
    untimed (device time of the call from HIP events on the context stream).
    Then, from where that run ends: a run with per-kernel timing on (force and
    drift launch averages), the steady step without rebuilds (the shortest of
    three 8-step calls started right after a rebuild) and a long window of at
    least five rebuild intervals."""
    ctx.set_state(r=state["r"], v=state["v"], F=state["F"])
    ctx.update_neighbor_lists()
    ctx.do_time_steps(args.warmup)
    reb0 = ctx.rebuild_count()
    t0 = time.perf_counter()
    ctx.do_time_steps(args.steps)  # synchronises on return
    wall = time.perf_counter() - t0
    ms = ctx.last_elapsed_ms()
    out = {"ms_per_step": ms / args.steps, "rebuilds": ctx.rebuild_count() - reb0, "wall_s": wall}
    ctx.set_kernel_timing(True)
    nt = max(args.steps, 200)
    ctx.do_time_steps(nt)
    kt = ctx.kernel_times()
    ctx.set_kernel_timing(False)
    out["force_kernel_avg_us"] = kt["force_ms"] / max(kt["force_launches"], 1) * 1e3
    out["drift_kernel_avg_us"] = kt["drift_ms"] / max(kt["drift_launches"], 1) * 1e3
    out["timed_run_launches"] = {"steps": nt, "force": kt["force_launches"], "drift": kt["drift_launches"]}
    steady = []
    for _ in range(3):
        ctx.update_neighbor_lists()
        r0 = ctx.rebuild_count()
        ctx.do_time_steps(8)
        if ctx.rebuild_count() == r0:
            steady.append(ctx.last_elapsed_ms() / 8)
    out["steady_ms_per_step"] = min(steady) if steady else None
    r0 = ctx.rebuild_count()
    ctx.do_time_steps(args.long_steps)
    out["long_window"] = {"steps": args.long_steps, "ms_per_step": ctx.last_elapsed_ms() / args.long_steps,
                          "rebuilds": ctx.rebuild_count() - r0}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(op_state, budget_s=15.0):
    """The oracle on the host, after the protocol of 10 untimed warm-up steps:
    the median rate of three runs of equal length (sized from the warm-up to
    about budget_s / 3 each), rebuilds included; once on one thread (the
    serial C restatement) and once on all host cores (OpenMP)."""
    import oracle
    r, v, F, D, L, N = op_state
    op = oracle.make_params(dims=3, N=N, box_size=L, force_cutoff=1.25, skin=0.3, D=D,
                            potential=oracle.POT_BERTHIER, system=oracle.SYS_NEWTONIAN, dt=0.005, kBT=0.2,
                            maxneighbours=128)
    st = oracle.State(op, r, v, D, F=F)
    st.update_neighbor_lists(cells=True)
    t0 = time.perf_counter()
    st.do_time_steps(10, cells=True)
    per = (time.perf_counter() - t0) / 10
    n = int(min(max(budget_s / 3 / max(per, 1e-6), 3), 2000))
    rates, reb = [], []
    for _ in range(3):
        b0 = int(st.rebuilds[0])
        t0 = time.perf_counter()
        st.do_time_steps(n, cells=True)
        rates.append(N * n / (time.perf_counter() - t0))
        reb.append(int(st.rebuilds[0]) - b0)
    model = cpu_model()
    out = {"value": float(np.median(rates)), "unit": "particle-steps/s", "cores": 1, "kind": "port",
           "cpu": model, "protocol": "10 warm-up steps, median of 3 runs",
           "sample": f"oracle (serial FP64 C restatement, cell-list rebuild) on the same N={N} state: "
                     f"3 runs of {n} steps ({reb} rebuilds), rates {[round(x) for x in rates]}"}
    threads = min(16, len(os.sched_getaffinity(0)))
    rb, vb, Fb = (np.ascontiguousarray(x, dtype=np.float64).copy() for x in (r, v, F))
    t0 = time.perf_counter()
    oracle.baseline_steps(op, 10, threads, rb, vb, Fb, D)
    per = (time.perf_counter() - t0) / 10
    n = int(min(max(budget_s / 3 / max(per, 1e-6), 3), 5000))
    rates = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.baseline_steps(op, n, threads, rb, vb, Fb, D)
        rates.append(N * n / (time.perf_counter() - t0))
    out["all_cores"] = {"value": float(np.median(rates)), "unit": "particle-steps/s", "cores": threads,
                        "kind": "port", "cpu": model, "protocol": "10 warm-up steps, median of 3 runs",
                        "sample": f"the same Newtonian step on {threads} OpenMP threads (per-particle rows, "
                                  f"gathered forces): 3 runs of {n} steps, rates {[round(x) for x in rates]}"}
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
    ctx, p = make_ctx(nat, N, L, D, nat.PREC_FP64, cell_divisor=args.cell_divisor, build_mode=args.build_mode)
    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else None
    if shape:
        ctx.set_build_launch(*shape)
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

    res64 = leg(ctx, state, args)
    ms64 = res64["ms_per_step"]
    log(f"[bench] fp64: {ms64 * 1e3:.1f} us/step ({res64['rebuilds']} rebuilds), steady "
        f"{res64['steady_ms_per_step'] * 1e3:.1f} us, long window {res64['long_window']['ms_per_step'] * 1e3:.1f} us")
    value = N / (ms64 / 1e3)

    f_avg_s = res64["force_kernel_avg_us"] / 1e6
    bytes_per_particle = 24 + 8 + 48 + 24 + 4 * (n_half + 1)
    achieved = N * bytes_per_particle / f_avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("traffic_k_force1_berthier3d_N1e6_fp64"),
                "kernel": "k_force<1> (pair forces + second half-kick + COM partials)",
                "kernel_avg_us": round(f_avg_s * 1e6, 2), "bytes_per_particle": round(bytes_per_particle, 2),
                "n_half_mean": round(n_half, 3), "drift_kernel_avg_us": round(res64["drift_kernel_avg_us"], 2)}
    if cc:
        roofline.update({"copy_ceiling_GBs": round(cc, 1), "frac_of_copy_ceiling": round(achieved / cc, 4)})
        log(f"[bench] measured copy ceiling {cc:.0f} GB/s")
    # whole-step bytes: the force kernel's plus the drift kernel's 184 B and the
    # rebuild's amortised share (8d+8 + 4 n_half + 8 per particle per rebuild)
    lw = res64["long_window"]
    step_bytes = bytes_per_particle + 184 + (32 + 4 * n_half + 8) * lw["rebuilds"] / lw["steps"]
    out = {"metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded jittered lattice, relaxed and thermalised on device)",
           "config": {"workload": f"3D polydisperse non-additive soft spheres, N={N}, rho=1.02, Newtonian NVE "
                                  f"dt=0.005, skin 0.3, kBT=0.2 start", "N": N, "dims": 3,
                      "potential": "Berthier (rc=1.25, zeta=0.2)", "dynamics": "Newtonian", "dt": 0.005},
           "roofline": roofline,
           "fp64": dict(res64, step_roofline_frac=round(N * step_bytes / (lw["ms_per_step"] / 1e3) / 1e9
                                                        / HBM_PEAK_GBS, 4))}

    if args.drift_steps > 0:
        ctx.set_state(r=state["r"], v=state["v"], F=state["F"])
        ctx.update_neighbor_lists()
        d64, fl64 = drift_run(ctx, args.drift_steps)
        out["fp64"].update({"nve_rel_drift_per_1e5_steps": d64, "nve_rel_fluct": fl64,
                            "drift_window_steps": args.drift_steps})

    if not args.no_mixed:
        ctxm, _ = make_ctx(nat, N, L, D, nat.PREC_MIXED, cell_divisor=args.cell_divisor, build_mode=args.build_mode)
        if shape:
            ctxm.set_build_launch(*shape)
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
        resm = leg(ctxm, state, args)
        msm = resm["ms_per_step"]
        out["mixed"] = dict(resm, value=N / (msm / 1e3), speedup_vs_fp64=ms64 / msm,
                            speedup_vs_fp64_steady=res64["steady_ms_per_step"] / resm["steady_ms_per_step"]
                            if resm["steady_ms_per_step"] else None,
                            speedup_vs_fp64_long_window=lw["ms_per_step"] / resm["long_window"]["ms_per_step"],
                            max_rel_force_err_vs_mean=float(dn.max() / fn.mean()),
                            max_rel_force_err_vs_max=float(dn.max() / fn.max()))
        if args.drift_steps > 0:
            ctxm.set_state(r=state["r"], v=state["v"], F=state["F"])
            ctxm.update_neighbor_lists()
            dm, flm = drift_run(ctxm, args.drift_steps)
            out["mixed"].update({"nve_rel_drift_per_1e5_steps": dm, "nve_rel_fluct": flm})
        log(f"[bench] mixed: {msm * 1e3:.1f} us/step ({resm['rebuilds']} rebuilds), steady "
            f"{resm['steady_ms_per_step'] * 1e3:.1f} us, long window {resm['long_window']['ms_per_step'] * 1e3:.1f} us, "
            f"force err {out['mixed']['max_rel_force_err_vs_mean']:.3g} of mean |F|")
        ctxm.close()
    else:
        ctx.close()

    small_states = None
    if not args.no_small:
        small, small_states = small_configs(args, nat)
        out["small_configs"] = small
    if not args.no_batched:
        out["batched"] = batched_summary(args, nat)
    if not args.no_served:
        from simulationcode_amd import serve
        out["served"] = serve.summary(args, log)
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, 1 thread and all host cores)")
        out["cpu_baseline"] = cpu_baseline((state["r"], state["v"], state["F"], D, L, N))
        if small_states is not None:
            out["cpu_baseline"]["small_configs"] = cpu_small_configs(small_states)
    else:
        out["cpu_baseline"] = None
    return out


def run_batched(nat, specs, cd, shape, streams, graph, warmup, steps, relax=(300, 0.005, 0.02), chunk=8):
    """Admit `specs` (cell divisor cd, build launch shape) into `streams`
    batches, relax, warm up, then time `steps` steps.  Returns the
    measurements."""
    if streams > 1:
        b = nat.BatchGroup(len(specs) + 1, streams=streams, graph=graph)
    else:
        b = nat.Batch(len(specs) + 1, graph=graph)
    b.set_build_launch(*shape)
    b.set_ahead(chunk)
    ids, Ns, dims = [], [], []
    t0 = time.perf_counter()
    for s in specs:
        p, D, r, v = W_replica(s, cd)
        ids.append(b.add(p, D, r, v))
        Ns.append(s["N"])
        dims.append(s["dims"])
    t_admit = time.perf_counter() - t0
    b.minimize(*relax)
    b.step(warmup)
    reb0 = sum(b.stats(i)["rebuilds"] for i in ids)
    b.step(steps)
    ms = b.info()["last_ms"]
    reb = sum(b.stats(i)["rebuilds"] for i in ids) - reb0
    tot = sum(Ns)
    # algorithmic bytes per particle-step, Brownian with on-device noise:
    # r 2*8d + F write 8d + disp 2*8d + D 8 + 4 (n_half + 1)
    byt = 0.0
    for rid, n, d in zip(ids, Ns, dims):
        t, _ = b.neighbor_stats(rid)
        byt += n * (40 * d + 8 + 4 * (t / 2.0 / n + 1))
    achieved = byt * steps / (ms / 1e3) / 1e9
    b.close() if hasattr(b, "close") else None
    return {"value": tot * steps / (ms / 1e3), "ms_per_step": ms / steps, "replicas": len(ids), "particles": tot,
            "rebuilds_per_step": reb / steps, "admit_s": round(t_admit, 2), "shape": shape,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "kernel": "whole batched step (all kernels)",
                         "bytes_per_particle_step_mean": round(byt / tot, 2)}}


def run_batched_stated_dt(nat, specs, cd, shape, steps=300, call=25, relax=(300, 0.005, 0.02)):
    """The batched workload at its stated dt = 0.001: one batch stepped in
    calls of `call` steps; replicas whose state goes non-finite (or whose
    lists overflow) are retired after the call, as the server does, and the
    others go on.  Reports the failures, the rebuild rate and the throughput
    of the steps made."""
    b = nat.Batch(len(specs) + 1)
    b.set_build_launch(*shape)
    ids, Ns = [], {}
    for s in specs:
        p, D, r, v = W_replica(dict(s, dt=0.001), cd)
        rid = b.add(p, D, r, v)
        ids.append(rid)
        Ns[rid] = s["N"]
    b.minimize(*relax)
    alive = set(ids)
    failed, done_steps, rebuilds = {}, {}, {}
    wall = 0.0
    made = 0
    while made < steps and alive:
        k = min(call, steps - made)
        t0 = time.perf_counter()
        try:
            b.step(k)
        except nat.ScError:
            pass
        wall += time.perf_counter() - t0
        made += k
        for rid in sorted(alive):
            code = b.replica_error(rid)
            if code != nat.SC_OK:
                st = b.stats(rid)
                failed[rid] = {"code": int(code), "steps": int(st["steps"])}
                rebuilds[rid] = int(st["rebuilds"])
                done_steps[rid] = int(st["steps"])
                b.remove(rid)
                alive.discard(rid)
    for rid in alive:
        st = b.stats(rid)
        done_steps[rid] = int(st["steps"])
        rebuilds[rid] = int(st["rebuilds"])
    b.close()
    psteps = sum(Ns[r] * done_steps[r] for r in ids)
    fsteps = sorted(f["steps"] for f in failed.values())
    mean_steps = sum(done_steps.values()) / len(ids)
    return {"dt": 0.001, "steps": steps, "call_steps": call, "replicas": len(ids),
            "failed": len(failed), "failed_frac": round(len(failed) / len(ids), 4),
            "failure_steps_p0_p50": [fsteps[0], fsteps[len(fsteps) // 2]] if fsteps else None,
            "rebuilds_per_step": round(sum(rebuilds.values()) / mean_steps, 3) if mean_steps else None,
            "rebuild_interval_steps_per_replica": round(mean_steps * len(ids) / max(sum(rebuilds.values()), 1), 2),
            "value": psteps / wall if wall else None, "unit": "particle-steps/s",
            "note": "steps made by every replica until the call after its failure; the stated dt with explicit "
                    "Euler-Maruyama on the r^-12 core"}


def W_replica(spec, cd):
    from simulationcode_amd import workloads as W
    return W.replica(**dict(spec, cell_divisor=cd))


def tuned_shape(nat, p, D, r, v, steps=20):
    """The build launch shape sc_autotune picks for a system, tuned on a
    copy of it (the measured context then starts from the stated state)."""
    ctx = nat.Context(p, D)
    ctx.set_state(r=r, v=v)
    ctx.update_neighbor_lists()
    ctx.calculate_new_forces(None)
    table, best = ctx.autotune(steps)
    ctx.close()
    return table[best]["block_size"], table[best]["lanes_per_particle"]


def fixed_shape(args):
    """The build launch shape given by --shape, or None."""
    return tuple(int(x) for x in args.shape.split(",")) if args.shape else None


def small_configs(args, nat):
    """The two small stated systems on the device (extra key `small_configs`),
    each from a seeded start, device time of the timed steps from HIP events:
    LJ: 3D Lennard-Jones N=1000, rho=1 (box 10), simple-cubic lattice with
    +-0.1 jitter, Newtonian velocity Verlet dt=0.01 from kBT=0.1, FP64, 1000
    steps.  Polydisperse: 3D soft spheres N=8000, rho=1.02, Langevin
    kBT=0.15, gamma=1, dt=0.01 (on-device noise) after 200 relaxation
    iterations (FP64), 10^4 steps in FP64 and in the mixed form from that
    state, with the mixed form's pair-force error there.  Both are
    latency-bound (a few microseconds of launch and dependency per kernel):
    the rate is what one small system gets alone.  The build launch shape of
    each is the one sc_autotune picks on a copy of the system (or --shape).  Returns the
    results and the states the CPU baseline times on."""
    from simulationcode_amd import workloads as W
    out, states = {}, {}
    p, D, r, v = W.replica(1000, 3, potential="lj", system="newtonian", kBT=0.1, dt=0.01, rho=1.0,
                           precision="fp64", seed=args.seed + 1, cell_divisor=0)
    shape = fixed_shape(args) or tuned_shape(nat, p, D, r, v)
    ctx = nat.Context(p, D)
    ctx.set_build_launch(*shape)
    ctx.set_state(r=r, v=v)
    ctx.update_neighbor_lists()
    ctx.calculate_new_forces(None)
    states["lj"] = (p, D, ctx.get_state())
    r0 = ctx.rebuild_count()
    ctx.do_time_steps(1000)
    ms = ctx.last_elapsed_ms()
    out["lj_N1000_newtonian_fp64"] = {"value": 1000 * 1000 / (ms / 1e3), "unit": "particle-steps/s",
                                      "us_per_step": ms, "steps": 1000, "rebuilds": ctx.rebuild_count() - r0,
                                      "build_shape": list(shape)}
    ctx.close()
    res = {}
    start = None
    for name, prec in (("fp64", nat.PREC_FP64), ("mixed", nat.PREC_MIXED)):
        p, D, r, v = W.replica(8000, 3, potential="berthier", system="langevin", kBT=0.15, dt=0.01, rho=1.02,
                               precision="mixed" if prec == nat.PREC_MIXED else "fp64", seed=args.seed + 2,
                               cell_divisor=0)
        ctx = nat.Context(p, D)
        if start is None:  # relax the seeded lattice once (FP64); both forms start from that state
            ctx.set_state(r=r, v=np.zeros_like(v))
            ctx.update_neighbor_lists()
            ctx.minimize(200, 0.005, 0.02)
            start = ctx.get_state(F=False)["r"]
            shape = fixed_shape(args) or tuned_shape(nat, p, D, start, v)
        ctx.set_build_launch(*shape)
        # pair forces on the common state: zero velocities and injected zero
        # noise, so that the Langevin terms vanish in both forms
        ctx.set_state(r=start, v=np.zeros_like(v))
        ctx.update_neighbor_lists()
        ctx.calculate_new_forces(np.zeros_like(v))
        Fpair = ctx.get_state(r=False, v=False, F=True)["F"]
        ctx.set_state(v=v)
        ctx.update_neighbor_lists()
        ctx.calculate_new_forces(None)
        if name == "fp64":
            states["poly"] = (p, D, ctx.get_state())
        r0 = ctx.rebuild_count()
        ctx.do_time_steps(10 ** 4)
        ms = ctx.last_elapsed_ms()
        res[name] = {"value": 8000 * 1e4 / (ms / 1e3), "unit": "particle-steps/s", "us_per_step": ms / 10.0,
                     "steps": 10 ** 4, "rebuilds": ctx.rebuild_count() - r0, "build_shape": list(shape),
                     "F": Fpair}
        ctx.close()
    F64, F32 = res["fp64"].pop("F"), res["mixed"].pop("F")
    fn = np.linalg.norm(F64, axis=1)
    res["mixed"].update(speedup_vs_fp64=res["mixed"]["value"] / res["fp64"]["value"],
                        max_rel_force_err_vs_mean=float(np.linalg.norm(F32 - F64, axis=1).max() / fn.mean()))
    out["poly_N8000_langevin"] = res
    log(f"[bench] small configs: LJ N=1000 {out['lj_N1000_newtonian_fp64']['us_per_step']:.1f} us/step; "
        f"polydisperse N=8000 Langevin FP64 {res['fp64']['us_per_step']:.1f} us/step, "
        f"mixed {res['mixed']['us_per_step']:.1f} us/step")
    return out, states


def cpu_small_configs(states, budget_s=4.0):
    """The oracle (serial FP64) on the same two small systems' starts: LJ
    N=1000 Newtonian, and the polydisperse N=8000 Langevin step with injected
    noise; steps sized to about budget_s / 2 each after 10 warm-up steps."""
    import oracle
    out = {}
    for key, name in (("lj", "lj_N1000_newtonian_fp64"), ("poly", "poly_N8000_langevin")):
        p, D, st = states[key]
        lj = key == "lj"
        op = oracle.make_params(dims=3, N=p.N, box_size=p.box_size, force_cutoff=p.force_cutoff,
                                skin=p.skin_distance, D=D, potential=oracle.POT_LJ if lj else oracle.POT_BERTHIER,
                                system=oracle.SYS_NEWTONIAN if lj else oracle.SYS_LANGEVIN, kBT=p.kBT,
                                gamma=p.gamma, dt=p.dt)
        ost = oracle.State(op, st["r"], st["v"], D, F=st["F"])
        ost.update_neighbor_lists(cells=True)
        rng = np.random.default_rng(7)
        noise = None if lj else rng.standard_normal((10, p.N, 3))
        t0 = time.perf_counter()
        ost.do_time_steps(10, noise, cells=True)
        per = (time.perf_counter() - t0) / 10
        n = int(min(max(budget_s / 2 / max(per, 1e-6), 10), 2000))
        noise = None if lj else rng.standard_normal((n, p.N, 3))
        t0 = time.perf_counter()
        ost.do_time_steps(n, noise, cells=True)
        dt_s = time.perf_counter() - t0
        out[name] = {"value": p.N * n / dt_s, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                     "sample": f"{n} steps after 10 warm-up steps" + ("" if lj else " (injected host noise)")}
    return out


# measured best launch choices of the 256-replica batch
BATCHED_CD, BATCHED_SHAPE, BATCHED_STREAMS, BATCHED_CHUNK = 2, (256, 0), 3, 12


def batched_summary(args, nat, steps=300, warmup=20):
    """The batched workload at a bounded length for the default run: the 256
    replicas at the best measured launch choices (cell divisor 2, the
    per-slot radius build {256, 0}, three streams, chunks of 12 steps)."""
    from simulationcode_amd import workloads as W
    t0 = time.perf_counter()
    specs = W.batched_specs(args.replicas)
    res = run_batched(nat, specs, BATCHED_CD, BATCHED_SHAPE, BATCHED_STREAMS, False, warmup, steps,
                      chunk=BATCHED_CHUNK)
    res.update({"steps": steps, "warmup": warmup, "dt": 1e-4, "dynamics": "Brownian", "dtype": "mixed",
                "cell_divisor": BATCHED_CD, "streams": BATCHED_STREAMS, "chunk": BATCHED_CHUNK,
                "wall_s": round(time.perf_counter() - t0, 1)})
    log(f"[bench] batched: {res['ms_per_step'] * 1e3:.1f} us/step, {res['value']:.3g} particle-steps/s, "
        f"roofline {res['roofline']['frac']:.3f}, {res['rebuilds_per_step']:.2f} rebuilds/step")
    res["stated_dt"] = sd = run_batched_stated_dt(nat, specs, BATCHED_CD, BATCHED_SHAPE)
    log(f"[bench] batched at dt=0.001: {sd['failed']}/{sd['replicas']} replicas non-finite within "
        f"{sd['steps']} steps (first at step {sd['failure_steps_p0_p50'][0] if sd['failed'] else '-'}), "
        f"{sd['rebuilds_per_step']} rebuilds/step, {sd['value']:.3g} particle-steps/s")
    return res


def bench_small(args, nat):
    """The stated small systems alone (the `small_configs` of the default
    run): LJ N=1000 Newtonian as the line's value, the N=8000 Langevin pair
    beside it."""
    small, _ = small_configs(args, nat)
    lj = small["lj_N1000_newtonian_fp64"]
    return {"metric": "particle-steps/s", "value": lj["value"], "unit": "particle-steps/s", "n_gpus": 1,
            "steps": lj["steps"], "warmup": 0, "ms_per_step": lj["us_per_step"] / 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded lattices)",
            "config": {"workload": "3D Lennard-Jones N=1000, rho=1, Newtonian dt=0.01 from kBT=0.1; 3D "
                                   "polydisperse N=8000, rho=1.02, Langevin kBT=0.15 dt=0.01 (FP64 and mixed)"},
            "small_configs": small, "roofline": None, "cpu_baseline": None}


def bench_batched(args, nat):
    from simulationcode_amd import workloads as W
    specs = W.batched_specs(args.replicas)
    t0 = time.perf_counter()
    if args.cell_divisor:
        shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else (256, 4)
        cd, tune = args.cell_divisor, None
    else:
        cd, shape, tune = W.autotune_group(specs, streams=args.streams, chunk=args.chunk)
    log(f"[bench] autotuner: cell divisor {cd}, build block {shape[0]} x {shape[1]} lanes/slot "
        f"({time.perf_counter() - t0:.1f}s)")
    res = run_batched(nat, specs, cd, shape, args.streams, args.graph, args.warmup, args.steps,
                      chunk=args.chunk)
    log(f"[bench] batched: {res['ms_per_step'] * 1e3:.1f} us/step over {res['particles']} particles")
    return {"metric": METRIC, "value": res["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "mixed (f32 pair / f64 state)",
            "data": "synthetic (seeded lattices relaxed on device)",
            "config": {"workload": f"{res['replicas']} independent polydisperse replicas, N log-uniform [500, 20000], "
                                   "75% 3D / 25% 2D, Brownian dt=1e-4, kBT U[0.1,0.3], per-replica seeds",
                       "replicas": res["replicas"], "particles": res["particles"], "graph": args.graph,
                       "streams": args.streams, "cell_divisor": cd, "build_block": res["shape"][0],
                       "build_lanes_per_slot": res["shape"][1], "chunk": args.chunk, "autotune": tune},
            "rebuilds_per_step": res["rebuilds_per_step"], "roofline": res["roofline"], "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--N", type=int, default=10 ** 6)
    ap.add_argument("--workload", choices=("large", "batched", "served", "small", "copycal"), default="large",
                    help="small: the stated small systems alone; copycal: five plain 256 MiB device copies, "
                         "the reference reading for HBM counters")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-copy-ceiling", action="store_true", help="large: skip the measured device-copy ceiling")
    ap.add_argument("--no-batched", action="store_true", help="large: skip the bounded batched summary")
    ap.add_argument("--no-served", action="store_true", help="large: skip the bounded served summary")
    ap.add_argument("--no-small", action="store_true", help="large: skip the stated small systems")
    ap.add_argument("--long-steps", type=int, default=1000,
                    help="large: steps of each form's long window (at least five rebuild intervals)")
    ap.add_argument("--graph", action="store_true",
                    help="batched: replay the step as a captured hipGraph (eager launches measured faster)")
    ap.add_argument("--streams", type=int, default=3,
                    help="batched: split the replicas over this many batches on their own streams")
    ap.add_argument("--drift-steps", type=int, default=20000)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261206)
    ap.add_argument("--cell-divisor", type=int, default=0,
                    help="0: library default (large) / autotuned (batched)")
    ap.add_argument("--shape", default="",
                    help="build launch shape BS,LPP (large: both contexts and the small systems; small: both "
                         "systems; batched: with --cell-divisor)")
    ap.add_argument("--build-mode", type=int, default=0, help="large: sc_params.build_mode (0 automatic)")
    ap.add_argument("--chunk", type=int, default=12,
                    help="batched: steps launched ahead per chunk (0: one launch sequence per step)")
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
    elif args.workload == "small":
        out = bench_small(args, nat)
    else:
        from simulationcode_amd import serve
        out = serve.bench(args, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
