"""Serving form: continuous batching of replica jobs on one device.

Jobs (a replica specification plus a step count) arrive over time.  The
server keeps one running replica batch (simulationcode_amd._native.Batch) and
loops:
  1. jobs whose arrival time has passed join the wait queue;
  2. waiting jobs are admitted in arrival order while a replica slot and the
     HBM budget (sc_replica_device_bytes) allow; admitted lattice starts are
     relaxed together in a staging batch (capped steepest descent), then enter
     the running batch with their own velocities;
  3. the running batch advances by a chunk = min(steps left over running
     jobs, max_chunk), so newly arrived work is admitted between chunks while
     other jobs keep their device-resident state;
  4. finished jobs are downloaded and retired, freeing slot and budget.
A job's trajectory depends only on its own specification (per-replica
Philox noise, per-replica reductions, chunk boundaries that do not change the
arithmetic), so its result is bitwise the same alone and in the mix.
"""
import collections
import time

import numpy as np

from . import _native as nat
from . import workloads as W

RELAX = (300, 0.005, 0.02)  # staging descent: iterations, dt, max step


class Job:
    def __init__(self, job_id, spec, steps, arrival):
        self.id = job_id
        self.spec = spec
        self.steps = int(steps)
        self.arrival = float(arrival)
        self.left = int(steps)
        self.rid = None
        self.admitted = None
        self.done = None
        self.result = None
        self.error = None
        self.particles = int(spec["N"])


class Server:
    def __init__(self, hbm_budget_bytes=64 << 30, max_replicas=512, max_chunk=1000, graph=False,
                 keep_results=False, clock=time.perf_counter, cell_divisor=0, build_shape=(256, 4)):
        self.budget = int(hbm_budget_bytes)
        self.max_replicas = int(max_replicas)
        self.max_chunk = int(max_chunk)
        self.cell_divisor = int(cell_divisor)  # applied to every job alike (it sets the slot order)
        self.batch = nat.Batch(self.max_replicas, graph=graph)
        self.batch.set_build_launch(*build_shape)  # launch shape: no effect on results
        self.keep_results = keep_results
        self.clock = clock
        self.used_bytes = 0
        self.peak_bytes = 0
        self.max_running = 0
        self.running = {}
        self.waiting = collections.deque()
        self.finished = []
        self.particle_steps = 0

    def _bytes(self, job):
        if not hasattr(job, "_bytes"):
            p, D, r, v = W.replica(**dict(job.spec, cell_divisor=self.cell_divisor))
            job._prep = (p, D, r, v)
            job._bytes = nat.replica_device_bytes(p, D)
        return job._bytes

    def _admit(self):
        group = []
        while self.waiting and len(self.running) + len(group) < self.max_replicas:
            job = self.waiting[0]
            need = self._bytes(job)
            if self.used_bytes + need > self.budget and (self.running or group):
                break
            self.waiting.popleft()
            self.used_bytes += need
            group.append(job)
        if not group:
            return
        staging = nat.Batch(len(group) + 1)
        sid = []
        for job in group:
            p, D, r, v = job._prep
            sid.append(staging.add(p, D, r, None))
        staging.minimize(*RELAX)
        now = self.clock()
        for job, s in zip(group, sid):
            p, D, r, v = job._prep
            r_relaxed = staging.get_state(s)["r"]
            job.rid = self.batch.add(p, D, r_relaxed, v)
            job.admitted = now
            self.running[job.rid] = job
            del job._prep
        staging.close()
        self.peak_bytes = max(self.peak_bytes, self.used_bytes)
        self.max_running = max(self.max_running, len(self.running))

    def _fail_broken(self):
        """Retire replicas whose state went non-finite or overflowed their
        lists; the others are independent and continue."""
        for rid in list(self.running):
            code = self.batch.replica_error(rid)
            if code != nat.SC_OK:
                job = self.running.pop(rid)
                self.batch.remove(rid)
                self.used_bytes -= job._bytes
                job.error = code
                job.left = 0
                job.done = self.clock()
                self.finished.append(job)

    def _retire(self):
        now = self.clock()
        for rid in [k for k, j in self.running.items() if j.left == 0]:
            job = self.running.pop(rid)
            if self.keep_results:
                job.result = self.batch.get_state(rid)
            self.batch.remove(rid)
            self.used_bytes -= job._bytes
            job.done = now
            self.finished.append(job)

    def run(self, jobs, t0=None):
        """Serve `jobs` (arrival times relative to t0) to completion."""
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        t0 = self.clock() if t0 is None else t0
        for j in pending:
            j.arrival += t0
        while pending or self.waiting or self.running:
            now = self.clock()
            while pending and pending[0].arrival <= now:
                self.waiting.append(pending.popleft())
            self._admit()
            if not self.running:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - self.clock(), 0.01)))
                continue
            k = min(min(j.left for j in self.running.values()), self.max_chunk)
            try:
                self.batch.step(k)
            except nat.ScError as e:
                if e.code not in (nat.SC_ERR_NONFINITE, nat.SC_ERR_OVERFLOW):
                    raise
                self._fail_broken()
            for j in self.running.values():
                j.left -= k
                self.particle_steps += j.particles * k
            self._retire()
        return self.finished

    def close(self):
        self.batch.close()


def run_alone(spec, steps, max_chunk=1000, cell_divisor=0):
    """A job served on its own (reference for the alone-vs-mix check)."""
    srv = Server(max_replicas=2, max_chunk=max_chunk, keep_results=True, cell_divisor=cell_divisor)
    job = Job(0, spec, steps, 0.0)
    srv.run([job])
    srv.close()
    return job.result


def make_jobs(n, rate, seed, **kw):
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        spec, steps = W.served_job(rng, k, **kw)
        jobs.append(Job(k, spec, steps, t))
    return jobs


def bench(args, log):
    """Sweep arrival rates; per rate report p50/p99 completion latency
    (arrival to result) and aggregate particle-steps/s over the makespan."""
    rates = args.rate or [2.0, 8.0]
    # autotuner on a pilot of the job stream's replicas (cell size, launch shape)
    rng = np.random.default_rng(args.seed)
    pilot = [W.served_job(rng, k)[0] for k in range(24)]
    t0 = time.perf_counter()
    cd, shape, _ = W.autotune_batch(pilot, pilot=24, steps=40)
    log(f"[serve] autotuner: cell divisor {cd}, build block {shape[0]} x {shape[1]} lanes/slot "
        f"({time.perf_counter() - t0:.1f}s)")
    sweep = []
    for lam in rates:
        jobs = make_jobs(args.jobs, lam, seed=args.seed + int(lam * 1000))
        srv = Server(max_chunk=1000, cell_divisor=cd, build_shape=shape)
        t0 = time.perf_counter()
        done = srv.run(jobs, t0=t0)
        makespan = time.perf_counter() - t0
        failed = sum(1 for j in done if j.error is not None)
        lat = np.array([j.done - j.arrival for j in done if j.error is None])
        rec = {"rate_jobs_per_s": lam, "jobs": len(done), "failed_jobs": failed, "p50_latency_s": float(np.percentile(lat, 50)),
               "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
               "particle_steps_per_s": srv.particle_steps / makespan, "makespan_s": makespan,
               "particle_steps": srv.particle_steps}
        log(f"[serve] rate {lam}: p50 {rec['p50_latency_s']:.2f}s p99 {rec['p99_latency_s']:.2f}s "
            f"{rec['particle_steps_per_s']:.3g} p-s/s")
        sweep.append(rec)
        srv.close()
    best = max(sweep, key=lambda r: r["particle_steps_per_s"])
    return {"metric": "particle-steps/s FP64 vs mixed @ stated force err; p50/p99 replica latency",
            "value": best["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1, "steps": None,
            "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "mixed (f32 pair / f64 state)", "data": "synthetic Poisson job stream",
            "config": {"workload": "served stream: Poisson arrivals, N log-uniform [1e3, 2e5], dims 2/3, "
                                   "Newtonian/Langevin/Brownian, steps log-uniform [1e3, 1e5]",
                       "jobs_per_rate": args.jobs, "cell_divisor": cd, "build_block": shape[0],
                       "build_lanes_per_slot": shape[1]},
            "p50_latency_s": best["p50_latency_s"], "p99_latency_s": best["p99_latency_s"], "sweep": sweep,
            "roofline": None, "cpu_baseline": None}
