"""GPU tests of the C-ABI's contract around the step loop.

* Per-kernel timing changes nothing: timed and untimed runs give bitwise the
  same trajectories, and the timed run counts one force and one drift launch
  per step across rebuilds.
* Errors: a device error stops the call, later stepping calls refuse until a
  rebuild clears it, and injected noise stays aligned with the device's step
  counter afterwards.
* The exported lists are the ones the last rebuild stored
  (src/NeighborLists.jl:82-104), not lists at the current positions.
* Failed-replica isolation: a replica that goes non-finite is reported alone,
  can be removed, and the others continue bitwise as they would alone; the
  server retires such a job and completes the others.
* The MD driver's log steps carry E_pot_test (src/MD.jl:623), and the
  minimiser's verbose report (src/InherentStructure.jl:57-69) leaves the
  descent unchanged.
* Diameters that are not finite and positive are refused by sc_set_state
  and sc_batch_add, and the state stays as it was.
* Positions outside the box are their periodic images: the lists and forces
  are those of the wrapped configuration.
"""
import re

import numpy as np
import pytest

import oracle
from simulationcode_amd import _native as nat, md, serve
from simulationcode_amd import workloads as W
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _ctx(op, rc, r, v, D, **kw):
    ctx = nat.Context(H.sc_params_from(op, rc, **kw), D)
    ctx.set_state(r=r, v=v)
    ctx.update_neighbor_lists()
    return ctx


@pytest.mark.parametrize("system", [oracle.SYS_NEWTONIAN, oracle.SYS_LANGEVIN, oracle.SYS_BROWNIAN])
def test_timed_and_untimed_steps_are_bitwise_equal(system):
    N, steps = 20000, 240
    dt = 0.0005 if system == oracle.SYS_BROWNIAN else 0.005
    op, r, D, v, rc = H.make_system("berthier", N, 3, seed=61, system=system, dt=dt, kBT=0.3)
    noise = None
    if system == oracle.SYS_LANGEVIN:
        noise = np.random.default_rng(5).standard_normal((steps, N, 3))
    out = []
    for timed in (False, True):
        ctx = _ctx(op, rc, r, v, D, cell_divisor=2)
        ctx.set_kernel_timing(timed)
        reb0 = ctx.rebuild_count()
        ctx.do_time_steps(steps // 2, None if noise is None else noise[: steps // 2])
        ctx.do_time_steps(steps - steps // 2, None if noise is None else noise[steps // 2:])
        kt = ctx.kernel_times()
        out.append((ctx.get_state(disp=True), ctx.rebuild_count() - reb0, kt))
        ctx.close()
    (a, ra, _), (b, rb, kt) = out
    for f in ("r", "v", "F", "disp"):
        assert np.array_equal(a[f], b[f]), f
    assert ra == rb and ra >= 3
    assert kt["force_launches"] == steps and kt["drift_launches"] == steps
    assert kt["force_ms"] > 0.0 and kt["drift_ms"] > 0.0


def test_error_refusal_and_noise_alignment_after_an_error():
    N, dims = 1000, 3
    op, r, D, v, rc = H.make_system("berthier", N, dims, seed=62, system=oracle.SYS_LANGEVIN, dt=0.005)
    ctx = _ctx(op, rc, r, v, D)
    ctx.do_time_steps(20)
    good = ctx.get_state()
    bad = good["v"].copy()
    bad[123, 0] = np.nan
    ctx.set_state(v=bad)
    rng = np.random.default_rng(9)
    with pytest.raises(nat.ScError) as e:
        ctx.do_time_steps(50, rng.standard_normal((50, N, dims)))
    assert e.value.code == nat.SC_ERR_NONFINITE
    # later stepping calls refuse with the recorded error
    with pytest.raises(nat.ScError) as e:
        ctx.do_time_steps(5, rng.standard_normal((5, N, dims)))
    assert e.value.code == nat.SC_ERR_NONFINITE
    with pytest.raises(nat.ScError) as e:
        ctx.calculate_new_forces(rng.standard_normal((N, dims)))
    assert e.value.code == nat.SC_ERR_NONFINITE
    # a valid state and a rebuild clear it; injected noise then lines up
    # with the device's step counter (checked against the oracle)
    ctx.set_state(r=good["r"], v=good["v"], F=good["F"])
    ctx.update_neighbor_lists()
    noise = rng.standard_normal((30, N, dims))
    ctx.do_time_steps(30, noise)
    st = oracle.State(op, good["r"], good["v"], D, F=good["F"])
    st.update_neighbor_lists()
    st.do_time_steps(30, noise)
    g = ctx.get_state()
    d = g["r"] - st.r
    d -= np.round(d / op.box_size) * op.box_size
    assert np.max(np.abs(d)) <= 1e-10
    assert np.max(np.abs(g["v"] - st.v)) <= 1e-10


def test_exported_lists_are_the_rebuild_time_lists():
    op, r, D, v, rc = H.make_system("berthier", 2000, 3, seed=63)
    ctx = _ctx(op, rc, r, v, D, cell_divisor=2)
    moved = r + np.random.default_rng(4).uniform(-0.12, 0.12, r.shape)
    moved -= np.floor(moved / op.box_size) * op.box_size
    ctx.set_state(r=moved)  # no rebuild: the stored lists stay those of r
    nf, full, nh, half = ctx.get_neighbor_lists()
    at_r = oracle.build_lists(op, r, D, cells=False)
    at_moved = oracle.build_lists(op, moved, D, cells=False)
    assert not np.array_equal(at_r.full_pairs(), at_moved.full_pairs())
    assert np.array_equal(nf, at_r.nfull)
    assert np.array_equal(oracle._pairs(nf, full - 1, both=True), at_r.full_pairs())
    assert np.array_equal(oracle._pairs(nh, half - 1, both=False), at_r.half_pairs())
    ctx.update_neighbor_lists()
    nf, full, nh, half = ctx.get_neighbor_lists()
    assert np.array_equal(oracle._pairs(nf, full - 1, both=True), at_moved.full_pairs())
    assert np.array_equal(oracle._pairs(nh, half - 1, both=False), at_moved.half_pairs())


ISO_SPECS = [
    dict(N=800, dims=3, system="newtonian", kBT=0.2, dt=0.005, seed=21, precision="fp64"),
    dict(N=1200, dims=3, system="langevin", kBT=0.2, dt=0.005, seed=22, precision="mixed"),
    dict(N=900, dims=2, system="brownian", kBT=0.15, dt=0.0002, seed=23, precision="mixed"),
    dict(N=1500, dims=2, system="langevin", kBT=0.25, dt=0.005, seed=24, precision="fp64"),
]
ISO_RELAX = (60, 0.005, 0.02)


def _alone(spec, chunks):
    b = nat.Batch(2)
    rid = b.add(*W.replica(**spec))
    b.minimize(*ISO_RELAX)
    for k in chunks:
        b.step(k)
    out = b.get_state(rid)
    b.close()
    return out


def test_failed_replica_is_isolated():
    b = nat.Batch(8)
    ids = [b.add(*W.replica(**s)) for s in ISO_SPECS]
    b.minimize(*ISO_RELAX)
    b.step(40)
    bad = b.get_state(ids[1])["v"].copy()
    bad[7, 2] = np.nan  # host data: the replica's next drift makes its positions non-finite
    b.set_state(ids[1], v=bad)
    with pytest.raises(nat.ScError) as e:
        b.step(60)
    assert e.value.code == nat.SC_ERR_NONFINITE
    codes = {rid: b.replica_error(rid) for rid in ids}
    assert codes[ids[1]] == nat.SC_ERR_NONFINITE
    assert all(codes[rid] == nat.SC_OK for rid in ids if rid != ids[1])
    b.remove(ids[1])
    b.step(50)
    for k, rid in enumerate(ids):
        if k == 1:
            continue
        got = b.get_state(rid)
        want = _alone(ISO_SPECS[k], (40, 60, 50))
        for f in ("r", "v", "F"):
            assert np.array_equal(got[f], want[f]), (k, f)
    b.close()


class _PlantingServer(serve.Server):
    """A server that, once job `victim` is admitted, writes a NaN into one of
    its velocities (host data through sc_batch_set_state): the job's next
    drift makes its positions non-finite."""

    victim = 1

    def _join(self, ready):
        super()._join(ready)
        for rid, job in self.running.items():
            if job.id == self.victim and not getattr(job, "planted", False):
                v = self.batch.get_state(rid)["v"].copy()
                v[3, 0] = np.nan
                self.batch.set_state(rid, v=v)
                job.planted = True


def test_server_retires_a_failed_job_and_completes_the_others():
    victim = dict(N=600, dims=3, system="newtonian", kBT=0.3, dt=0.005, seed=31, precision="fp64")
    specs = [(ISO_SPECS[0], 300), (victim, 300), (ISO_SPECS[3], 250), (ISO_SPECS[2], 200)]
    jobs = [serve.Job(k, s, n, arrival=0.005 * k) for k, (s, n) in enumerate(specs)]
    srv = _PlantingServer(max_chunk=100, keep_results=True)
    done = srv.run(jobs)
    srv.close()
    by_id = {j.id: j for j in done}
    assert len(by_id) == len(specs)
    assert by_id[1].error == nat.SC_ERR_NONFINITE and by_id[1].result is None
    for k in (0, 2, 3):
        j = by_id[k]
        assert j.error is None
        alone = serve.run_alone(j.spec, j.steps, max_chunk=100)
        for f in ("r", "v", "F"):
            assert np.array_equal(j.result[f], alone[f]), (k, f)


def test_md_driver_logs_e_pot_test():
    op, r, D, v, rc = H.make_system("berthier", 1000, 3, seed=64)
    N, dims, L = op.N, op.dims, op.box_size
    S = md.Newtonian(0.2, 0.005, 1.0, False, dims)
    U = md.Berthier(rc=1.25)
    P = md.ParameterStruct(100, 1, N / L ** dims, N, L, 2000, 0.0, 0.1, 1.25 ** 2, 0.0, S, U, None, None)
    A = md.ArrayStruct(N, dims)
    A.r_array[...] = r
    A.D_array[...] = D
    out = md.OutputStruct()
    nl = md.initialize_neighbor_struct(0.3, L, 1.25, N, D, dims)
    lines = []
    md.perform_molecular_dynamics(A, P, out, nl, log=lines.append)
    assert len(lines) == 3  # log steps 1000 and 2000 (src/AuxiliaryFunctions.jl:637-646), then the summary
    assert all("E_pot_test = " in ln for ln in lines[:-1])
    want = oracle.energy_no_neigh(op, A.r_array, D)
    assert abs(out.E_pot_test - want) <= 1e-12 * abs(want)
    assert abs(out.E_pot_test - out.potential_energy) <= 1e-11 * abs(want)
    lines2 = []
    A.r_array[...] = r
    md.perform_molecular_dynamics(A, P, md.OutputStruct(), md.initialize_neighbor_struct(0.3, L, 1.25, N, D, dims),
                                  log=lines2.append, energy_test=False)
    assert not any("E_pot_test" in ln for ln in lines2)


def test_minimiser_verbose_report_leaves_the_descent_unchanged(capsys):
    op, r, D, v, rc = H.make_system("berthier", 343, 3, seed=65, relax=False)
    N, dims, L = op.N, op.dims, op.box_size

    def objects():
        S = md.Newtonian(0.2, 0.005, 1.0, False, dims)
        P = md.ParameterStruct(100, 1, N / L ** dims, N, L, 10, 0.0, 0.1, 1.25 ** 2, 0.0, S, md.Berthier(rc=1.25),
                               None, None)
        A = md.ArrayStruct(N, dims)
        A.r_array[...] = r
        A.D_array[...] = D
        return A, P, md.OutputStruct(), md.initialize_neighbor_struct(0.3, L, 1.25, N, D, dims)

    A, P, out, nl = objects()
    r1 = md.find_inherent_structure(r, P, A, nl, out, tolerance=1e-6, verbose=1)
    txt1 = capsys.readouterr().out
    A2, P2, out2, nl2 = objects()
    r2 = md.find_inherent_structure(r, P2, A2, nl2, out2, tolerance=1e-6, verbose=2)
    txt2 = capsys.readouterr().out
    assert np.array_equal(r1, r2)
    assert out.N_neighbor_list_rebuilds == out2.N_neighbor_list_rebuilds
    it = int(re.search(r"after (\d+) iterations", txt1).group(1))
    assert f"after {it} iterations" in txt2
    reports = [ln for ln in txt2.splitlines() if ln.startswith("(iteration, maxchange")]
    assert it > 1000 and len(reports) == 1 + (it - 1) // 1000
    assert reports[0].split("= (")[1].startswith("1, ")


@pytest.mark.parametrize("bad", [np.nan, 0.0, -0.5])
def test_bad_diameters_are_refused_and_the_state_kept(bad):
    op, r, D, v, rc = H.make_system("berthier", 500, 3, seed=9)
    ctx = _ctx(op, rc, r, v, D)
    ctx.calculate_new_forces(None)
    before = ctx.get_state()
    D2 = D.copy()
    D2[7] = bad
    with pytest.raises(nat.ScError) as e:
        ctx.set_state(r=r + 0.01, D=D2)
    assert e.value.code == nat.SC_ERR_ARG
    after = ctx.get_state()
    for f in ("r", "v", "F"):
        assert np.array_equal(before[f], after[f]), f
    ctx.do_time_steps(20)  # still usable
    assert np.all(np.isfinite(ctx.get_state()["r"]))
    ctx.close()
    b = nat.Batch(2)
    with pytest.raises(nat.ScError) as e:
        b.add(H.sc_params_from(op, rc), D2, r, v)
    assert e.value.code == nat.SC_ERR_ARG
    assert b.info()["active"] == 0
    b.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_positions_outside_the_box_are_their_periodic_images(dims):
    """The reference takes positions as given and works with them modulo the
    box (minimum images in the force loops, src/MD.jl:107-133; CellListMap's
    periodic cells in update_neighbor_lists!).  Moving a tenth of the
    particles by +-L leaves the pair sets and the forces as they were."""
    op, r, D, v, rc = H.make_system("berthier", 900 if dims == 2 else 1000, dims, seed=21)
    L = op.box_size
    rng = np.random.default_rng(3)
    shifted = r.copy()
    pick = rng.choice(len(r), len(r) // 10, replace=False)
    shifted[pick] += rng.choice([-L, L], size=(len(pick), dims))
    out = []
    for pos in (r, shifted):
        for mode, shape in ((0, (256, 0)), (0, (256, 1)), (2, None)):
            ctx = nat.Context(H.sc_params_from(op, rc, cell_divisor=2, build_mode=mode), D)
            if shape:
                ctx.set_build_launch(*shape)
            ctx.set_state(r=pos, v=v)
            ctx.update_neighbor_lists()
            ctx.calculate_new_forces(None)
            nf, full, nh, half = ctx.get_neighbor_lists()
            out.append((oracle._pairs(nf, full - 1, both=True), oracle._pairs(nh, half - 1, both=False),
                        ctx.get_state()["F"]))
            ctx.close()
    lst = oracle.build_lists(op, r, D, cells=False)
    for fp, hp, F in out:
        assert np.array_equal(fp, lst.full_pairs())
        assert np.array_equal(hp, lst.half_pairs())
        assert H.force_rel_err(F, out[0][2]) <= 1e-12
