"""GPU tests of the replica batch: every replica's trajectory is bitwise the
same alone and in a mix (and the same as a single-system context with the same
seed), hipGraph replay equals eager launches, and retiring a replica does not
disturb the others."""
import numpy as np
import pytest

from simulationcode_amd import _native as nat
from simulationcode_amd import workloads as W

pytestmark = pytest.mark.gpu

SPECS = [
    dict(N=700, dims=3, system="newtonian", kBT=0.2, dt=0.005, seed=1, precision="fp64"),
    dict(N=900, dims=2, system="brownian", kBT=0.15, dt=0.001, seed=2, precision="mixed"),
    dict(N=1500, dims=3, system="langevin", kBT=0.3, dt=0.005, seed=3, precision="mixed"),
    dict(N=512, dims=3, system="brownian", potential="lj", kBT=0.2, dt=0.0002, seed=4, precision="fp64"),
    dict(N=2000, dims=2, system="langevin", kBT=0.1, dt=0.005, seed=5, precision="fp64"),
]
RELAX = (60, 0.005, 0.02)


def run(specs, steps, graph=False, chunks=1):
    b = nat.Batch(16, graph=graph)
    ids = []
    for s in specs:
        p, D, r, v = W.replica(**s)
        ids.append(b.add(p, D, r, v))
    b.minimize(*RELAX)
    for _ in range(chunks):
        b.step(steps // chunks)
    out = [b.get_state(i) for i in ids]
    stats = [b.stats(i) for i in ids]
    b.close()
    return out, stats


def test_same_alone_and_in_mix():
    mixed, mstats = run(SPECS, 150)
    for k, s in enumerate(SPECS):
        alone, astats = run([s], 150)
        for f in ("r", "v", "F"):
            assert np.array_equal(mixed[k][f], alone[0][f]), (k, f)
        assert mstats[k]["rebuilds"] == astats[0]["rebuilds"]
        assert np.all(np.isfinite(mixed[k]["r"]))


def test_batch_equals_single_context():
    for s in (SPECS[0], SPECS[1]):
        p, D, r, v = W.replica(**s)
        ctx = nat.Context(p, D)
        ctx.set_state(r=r, v=v)
        ctx.update_neighbor_lists()
        ctx.minimize(RELAX[0], RELAX[1], RELAX[2])
        ctx.do_time_steps(120)
        g = ctx.get_state()
        b_out, _ = run([s], 120)
        for f in ("r", "v", "F"):
            assert np.array_equal(g[f], b_out[0][f]), f


def test_graph_replay_matches_eager():
    be = nat.Batch(16)
    bg = nat.Batch(16, graph=True)
    ids_e, ids_g = [], []
    for s in SPECS:
        p, D, r, v = W.replica(**s)
        ids_e.append(be.add(p, D, r, v))
        ids_g.append(bg.add(p, D, r, v))
    be.minimize(*RELAX)
    bg.minimize(*RELAX)
    for rep in range(4):  # several replays, eager host work in between
        be.step(25)
        bg.step(25)
        for a, c in zip(ids_e, ids_g):
            se, sg = be.get_state(a), bg.get_state(c)
            for f in ("r", "v", "F"):
                assert np.array_equal(se[f], sg[f]), (rep, f)
    assert bg.stats(ids_g[0])["steps"] == 100


def test_remove_does_not_disturb_others():
    b = nat.Batch(16)
    ids = []
    for s in SPECS[:3]:
        p, D, r, v = W.replica(**s)
        ids.append(b.add(p, D, r, v))
    b.minimize(*RELAX)
    b.step(50)
    b.remove(ids[1])
    b.step(50)
    got = [b.get_state(ids[0]), b.get_state(ids[2])]
    ref0, _ = run([SPECS[0]], 100, chunks=2)
    ref2, _ = run([SPECS[2]], 100, chunks=2)
    assert np.array_equal(got[0]["r"], ref0[0]["r"])
    assert np.array_equal(got[1]["r"], ref2[0]["r"])
    assert b.info()["active"] == 2


def test_admission_bytes_and_info():
    p, D, r, v = W.replica(**SPECS[2])
    need = nat.replica_device_bytes(p, D)
    assert need > p.N * 200 * 4
    b = nat.Batch(4)
    b.add(p, D, r, v)
    info = b.info()
    assert info["active"] == 1 and info["particles"] == p.N and info["device_bytes"] == need


def test_stream_group_and_async_match_single_batch():
    ref, _ = run(SPECS, 80)
    g = nat.BatchGroup(16, streams=2, graph=True)
    ids = []
    for s in SPECS:
        p, D, r, v = W.replica(**s)
        ids.append(g.add(p, D, r, v))
    g.minimize(*RELAX)
    g.step(30)
    g.step(50)
    assert g.info()["active"] == len(SPECS)
    for k, gid in enumerate(ids):
        got = g.get_state(gid)
        for f in ("r", "v", "F"):
            assert np.array_equal(got[f], ref[k][f]), (k, f)
    g.close()
    # a call made while steps are in flight waits for them first
    b = nat.Batch(4)
    p, D, r, v = W.replica(**SPECS[0])
    rid = b.add(p, D, r, v)
    b.minimize(*RELAX)
    b.step_async(80)
    st = b.stats(rid)
    assert st["steps"] == 80
    assert np.array_equal(b.get_state(rid)["r"], ref[0]["r"])
    b.close()


def test_chunked_ahead_equals_eager_stepping():
    """Steps enqueued ahead in chunks (replicas halting on their own skin
    triggers, rebuilt together after each chunk) give bitwise the results of
    one launch sequence per step, for every replica of a mix of dynamics,
    dimensions and precisions, over call lengths that end mid-chunk; rebuild
    counts agree."""
    specs = SPECS + [dict(N=3000, dims=3, system="brownian", kBT=0.3, dt=0.0005, seed=6, precision="mixed"),
                     dict(N=1200, dims=2, system="newtonian", kBT=0.3, dt=0.006, seed=7, precision="fp64")]
    out = []
    for ahead in (False, True):
        b = nat.Batch(16)
        b.set_ahead(ahead)
        ids = [b.add(*W.replica(**s)) for s in specs]
        b.minimize(*RELAX)
        for k in (37, 1, 100, 16, 201):
            b.step(k)
        out.append(([b.get_state(i) for i in ids], [b.stats(i) for i in ids]))
        b.close()
    (se, te), (sa, ta) = out
    for k in range(len(specs)):
        for f in ("r", "v", "F"):
            assert np.array_equal(se[k][f], sa[k][f]), (k, f)
        assert te[k]["rebuilds"] == ta[k]["rebuilds"] and te[k]["steps"] == ta[k]["steps"] == 355
    assert sum(t["rebuilds"] for t in ta) > 3 * len(specs)


def test_admission_during_an_async_step_waits_for_it():
    """sc_batch_add while a chunked asynchronous step is in flight settles
    that step first: the replicas end bitwise as with a synchronous step
    before the admission, and the admitted replica steps as it would alone."""
    first, late = SPECS[:3], SPECS[3]
    out = []
    for asynchronous in (False, True):
        b = nat.Batch(16)
        ids = [b.add(*W.replica(**s)) for s in first]
        b.minimize(*RELAX)
        if asynchronous:
            b.step_async(90)
        else:
            b.step(90)
        p, D, r, v = W.replica(**late)
        ids.append(b.add(p, D, r, v))
        b.minimize(*RELAX)  # the admitted lattice start (the same in both orders)
        b.step(40)
        out.append(([b.get_state(i) for i in ids], [b.stats(i)["steps"] for i in ids]))
        b.close()
    (sa, na), (sb, nb) = out
    assert na == nb == [130, 130, 130, 40]
    for k in range(len(sa)):
        for f in ("r", "v", "F"):
            assert np.array_equal(sa[k][f], sb[k][f]), (k, f)
