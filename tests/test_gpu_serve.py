"""GPU test of the serving loop: jobs admitted while others run (HBM budget
forces queueing, chunks interleave) give bitwise the same results as each job
served alone."""
import numpy as np
import pytest

from simulationcode_amd import _native as nat
from simulationcode_amd import serve
from simulationcode_amd import workloads as W

pytestmark = pytest.mark.gpu

SPECS = [
    (dict(N=800, dims=3, system="newtonian", kBT=0.2, dt=0.005, seed=11, precision="mixed"), 300),
    (dict(N=1200, dims=2, system="brownian", kBT=0.15, dt=0.001, seed=12, precision="mixed"), 500),
    (dict(N=600, dims=3, system="langevin", kBT=0.25, dt=0.005, seed=13, precision="fp64"), 250),
    (dict(N=2000, dims=3, system="brownian", kBT=0.1, dt=0.001, seed=14, precision="mixed"), 400),
    (dict(N=900, dims=2, system="newtonian", kBT=0.3, dt=0.005, seed=15, precision="fp64"), 350),
]


def test_served_results_equal_alone():
    sizes = []
    for spec, _ in SPECS:
        p, D, r, v = W.replica(**spec)
        sizes.append(nat.replica_device_bytes(p, D))
    budget = sizes[0] + sizes[1] + sizes[2]
    jobs = [serve.Job(k, spec, steps, arrival=0.01 * k) for k, (spec, steps) in enumerate(SPECS)]
    srv = serve.Server(hbm_budget_bytes=budget, max_chunk=120, keep_results=True)
    done = srv.run(jobs)
    srv.close()
    assert len(done) == len(SPECS)
    assert srv.max_running >= 2          # jobs overlapped on the device
    assert srv.peak_bytes <= budget
    for j in done:
        alone = serve.run_alone(j.spec, j.steps)
        for f in ("r", "v", "F"):
            assert np.array_equal(j.result[f], alone[f]), (j.id, f)
        assert j.done >= j.admitted >= j.arrival
