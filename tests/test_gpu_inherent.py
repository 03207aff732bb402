"""GPU parity of the inherent-structure minimiser (find_inherent_structure,
src/InherentStructure.jl:4-87) against the oracle on identical starts.

Bars: identical iteration counts and rebuild counts; positions and
displacements within 1e-9 (absolute; box sides 5-12); the last iteration's
forces within 1e-11 absolute per particle (near a minimum they are sums of
O(1) pair forces cancelling to |F| ~ tolerance, so a bar relative to max |F|
would measure the cancellation, not the kernel); v (the last capped step)
within 1e-12."""
import numpy as np
import pytest

import oracle
from simulationcode_amd import _native as nat, md
from tests import helpers as H

pytestmark = pytest.mark.gpu

DT, MAX_STEP = 0.005, 0.001


def _ctx(op, rc, r, v, D):
    ctx = nat.Context(H.sc_params_from(op, rc, nat.PREC_FP64, 1, 1234, 0), D)
    ctx.set_state(r=r, v=v)
    ctx.update_neighbor_lists()
    return ctx


def _wrapped_diff(a, b, L):
    d = a - b
    return d - np.round(d / L) * L


@pytest.mark.parametrize("dims,N,tol,max_it", [(3, 125, 1e-4, 100000), (2, 144, 1e-4, 100000),
                                               (3, 256, 1e-3, 100000), (3, 125, 1e-12, 3000)])
def test_minimize_until_matches_oracle(dims, N, tol, max_it):
    op, r, D, v, rc = H.make_system("berthier", N, dims, seed=31, relax=False)
    st = oracle.State(op, r, v, D)
    it_o, fm_o, ok_o = st.find_inherent_structure(tolerance=tol, max_iterations=max_it, max_step=MAX_STEP, dt=DT)
    ctx = _ctx(op, rc, r, v, D)
    print(f"minimize_until dims={dims} N={N} tol={tol} oracle iterations={it_o}", flush=True)
    it_g, fm_g, ok_g = ctx.minimize_until(max_it, DT, MAX_STEP, tol)
    assert ok_g == ok_o and ok_o == (tol > 1e-6)
    assert it_g == it_o
    assert fm_g == pytest.approx(fm_o, rel=1e-10, abs=1e-13)
    assert ctx.rebuild_count() == int(st.rebuilds[0])
    g = ctx.get_state(r=True, v=True, F=True, disp=True)
    assert np.max(np.abs(_wrapped_diff(g["r"], st.r, op.box_size))) <= 1e-9
    assert np.max(np.abs(g["disp"] - st.disp)) <= 1e-9
    assert np.max(np.linalg.norm(g["F"] - st.F, axis=1)) <= 1e-11
    assert np.max(np.abs(g["v"] - st.v)) <= 1e-12


def test_minimize_until_twice_and_then_fixed_iterations():
    """A converged state re-run converges at once (its max |F| is already at
    the tolerance); fixed-count sc_minimize still moves particles after a
    tolerance run (the stop state is cleared)."""
    op, r, D, v, rc = H.make_system("berthier", 125, 3, seed=31, relax=False)
    ctx = _ctx(op, rc, r, v, D)
    it1, fm1, ok1 = ctx.minimize_until(100000, DT, MAX_STEP, 1e-3)
    assert ok1 and it1 > 1
    it2, fm2, ok2 = ctx.minimize_until(100000, DT, MAX_STEP, 1e-3)
    assert ok2 and it2 <= 3 and fm2 <= 1e-3
    r0 = ctx.get_state(r=True)["r"]
    fm3 = ctx.minimize(50, DT, MAX_STEP)
    assert fm3 <= 1e-3
    r1 = ctx.get_state(r=True)["r"]
    assert np.max(np.abs(_wrapped_diff(r1, r0, op.box_size))) > 0.0


def test_minimize_until_rejects_bad_arguments():
    op, r, D, v, rc = H.make_system("berthier", 125, 3, seed=31, relax=False)
    ctx = _ctx(op, rc, r, v, D)
    for args in [(-1, DT, MAX_STEP, 1e-4), (10, 0.0, MAX_STEP, 1e-4), (10, DT, 0.0, 1e-4), (10, DT, MAX_STEP, -1.0)]:
        with pytest.raises(nat.ScError):
            ctx.minimize_until(*args)


def _md_objects(r, D, L, dims):
    N = len(D)
    S = md.Newtonian(0.2, 0.005, 1.0, False, dims)
    U = md.Berthier(rc=1.25)
    P = md.ParameterStruct(100, 1, N / L ** dims, N, L, 10, 0.0, 0.1, 1.25 ** 2, 1.5, S, U, None, None)
    A = md.ArrayStruct(N, dims)
    A.r_array[...] = r
    A.D_array[...] = D
    return A, P, md.OutputStruct(), md.initialize_neighbor_struct(0.3, L, 1.25, N, D, dims)


def test_find_inherent_structure_mirror(capsys):
    """The Python mirror: same positions as the oracle, potential energies
    recomputed, the reference's closing line printed, rebuilds counted, and
    "Not converged" raised past max_iterations."""
    op, r, D, v, rc = H.make_system("berthier", 125, 3, seed=31, relax=False)
    st = oracle.State(op, r, v, D)
    it_o, _, ok_o = st.find_inherent_structure(tolerance=1e-4, max_iterations=100000, max_step=MAX_STEP, dt=DT)
    assert ok_o
    A, P, out, nl = _md_objects(r, D, op.box_size, 3)
    rmin = md.find_inherent_structure(r, P, A, nl, out, tolerance=1e-4)
    assert np.max(np.abs(_wrapped_diff(rmin, st.r, op.box_size))) <= 1e-9
    assert out.N_neighbor_list_rebuilds == int(st.rebuilds[0])
    lists = oracle.build_lists(op, st.r, D)
    np.testing.assert_allclose(A.potential_energy_array, oracle.potential_energies(op, st.r, D, lists),
                               rtol=1e-9, atol=1e-12)
    assert f"after {it_o} iterations" in capsys.readouterr().out
    A, P, out, nl = _md_objects(r, D, op.box_size, 3)
    with pytest.raises(RuntimeError, match="Not converged"):
        md.find_inherent_structure(r, P, A, nl, out, tolerance=1e-4, max_iterations=50, verbose=0)


def test_batch_minimize_until_equals_each_replica_alone():
    """sc_batch_minimize_until: every replica stops after the same number of
    iterations, at bitwise the same positions and forces, as a single-system
    context running sc_minimize_until on it alone; replicas converge at
    different iterations, and replicas of the mix keep being stepped
    correctly afterwards."""
    from simulationcode_amd import workloads as W
    specs = [
        dict(N=700, dims=3, system="newtonian", kBT=0.2, dt=0.005, seed=1, precision="fp64"),
        dict(N=900, dims=2, system="brownian", kBT=0.15, dt=0.001, seed=2, precision="mixed"),
        dict(N=1500, dims=3, system="langevin", kBT=0.3, dt=0.005, seed=3, precision="fp64"),
        dict(N=400, dims=3, system="newtonian", kBT=0.2, dt=0.005, seed=6, precision="fp64"),
    ]
    tol, max_it = 1e-3, 20000
    b = nat.Batch(8, graph=True)
    ids = []
    for s in specs:
        p, D, r, v = W.replica(**s)
        ids.append(b.add(p, D, r, v))
    b.minimize(60, 0.005, 0.02)
    b.step(3)  # a captured step exists before the descent
    res = b.minimize_until(max_it, DT, MAX_STEP, tol)
    its = []
    for k, s in enumerate(specs):
        p, D, r, v = W.replica(**s)
        ctx = nat.Context(p, D)
        ctx.set_state(r=r, v=v)
        ctx.update_neighbor_lists()
        ctx.minimize(60, 0.005, 0.02)
        ctx.do_time_steps(3)
        it, fm, ok = ctx.minimize_until(max_it, DT, MAX_STEP, tol)
        assert res[ids[k]] == (it, ok), (k, res[ids[k]], it, ok)
        its.append(it)
        g = ctx.get_state()
        gb = b.get_state(ids[k])
        for f in ("r", "F"):
            assert np.array_equal(g[f], gb[f]), (k, f)
        # stepping continues identically after the descent (the noise
        # streams are keyed by the count of force evaluations)
        ctx.do_time_steps(20)
        specs[k]["_r"] = ctx.get_state()["r"]
    assert len(set(its)) > 1
    b.step(20)
    for k in range(len(specs)):
        assert np.array_equal(b.get_state(ids[k])["r"], specs[k]["_r"]), k
    b.close()


def test_minimize_until_edge_cases():
    """max_iters = 0 does nothing and reports not converged (the reference
    raises before its first iteration); tolerance 0 is never met, every
    iteration runs; an empty batch returns no results."""
    op, r, D, v, rc = H.make_system("berthier", 125, 3, seed=31, relax=False)
    ctx = _ctx(op, rc, r, v, D)
    r0 = ctx.get_state(r=True)["r"]
    it, fm, ok = ctx.minimize_until(0, DT, MAX_STEP, 1e-4)
    assert it == 0 and not ok
    assert np.array_equal(ctx.get_state(r=True)["r"], r0)
    it, fm, ok = ctx.minimize_until(40, DT, MAX_STEP, 0.0)
    assert it == 40 and not ok and fm > 0.0
    b = nat.Batch(4)
    assert b.minimize_until(10, DT, MAX_STEP, 1e-4) == {}
    from simulationcode_amd import workloads as W
    p, Dr, rr, vr = W.replica(N=300, dims=2, system="newtonian", kBT=0.2, dt=0.005, seed=9, precision="fp64")
    rid = b.add(p, Dr, rr, vr)
    assert b.minimize_until(25, DT, MAX_STEP, 0.0) == {rid: (25, False)}
    b.close()
