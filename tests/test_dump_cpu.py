"""Dump directory and restart on the host (no GPU): the datasets and
attributes that save_data / prepare_savefile write read back through
read_restart (src/Dump.jl:60-156, src/Restart.jl:1-104)."""
import numpy as np
import pytest

from simulationcode_amd import md

SYSTEMS = {
    "newtonian": lambda d: md.Newtonian(0.2, 0.005, 1.0, True, d),
    "langevin": lambda d: md.Langevin(0.15, 1.0, 0.01, 1.0, d),
    "brownian": lambda d: md.Brownian(0.2, 1.0, 1e-4, d),
}


@pytest.mark.parametrize("pot", ["berthier", "lj"])
@pytest.mark.parametrize("system", sorted(SYSTEMS))
def test_dump_round_trip(tmp_path, pot, system):
    N, dims = 50, 3 if pot == "berthier" else 2
    rng = np.random.default_rng(5)
    S = SYSTEMS[system](dims)
    U = md.Berthier(rc=1.25, ζ=0.2) if pot == "berthier" else md.LJ(1.0, 1.0, 2.5)
    fn = str(tmp_path / "run")
    di = md.DumpInfo(True, fn, {0, 10, 20}, True, True, True, True, True)
    P = md.ParameterStruct(100, 7, 1.0, N, 9.5, 20, 0.0, 0.1, 1.25 ** 2, 0.0, S, U, di, None)
    A = md.ArrayStruct(N, dims)
    A.r_array[...] = rng.uniform(0, 9.5, (N, dims))
    A.v_array[...] = rng.standard_normal((N, dims))
    A.F_array[...] = rng.standard_normal((N, dims))
    A.D_array[...] = rng.uniform(0.8, 1.2, N)
    nl = md.initialize_neighbor_struct(0.3, 9.5, 1.25, N, A.D_array, dims, maxneighbours=64)
    md.prepare_savefile(P, A, nl)
    w = md.dump_writer(fn)
    for name, arr in (("positions", A.r_array), ("velocities", A.v_array), ("forces", A.F_array)):
        w(10, name, arr)
    # the stored layout is the reference's dims x N
    assert np.load(f"{fn}/positions/10.npy").shape == (dims, N)
    A2, P2, out2, nl2 = md.read_restart(fn, 10)
    for f in ("r_array", "v_array", "F_array", "D_array"):
        assert np.array_equal(getattr(A, f), getattr(A2, f)), f
    assert out2.steps_done == 10
    assert type(P2.system) is type(S) and type(P2.interaction_potential) is type(U)
    for k, v in vars(S).items():
        assert getattr(P2.system, k) == v, k
    for k, v in vars(U).items():
        assert getattr(P2.interaction_potential, k) == pytest.approx(v, rel=1e-15), k
    for k in ("N", "box_size", "N_steps", "force_cutoff2", "q_cutoff", "random_seed", "ρ"):
        assert getattr(P2, k) == getattr(P, k), k
    assert P2.dump_info.when_to_save_array == {0, 10, 20}
    assert nl2.skin_distance == 0.3 and nl2.maxneighbours == 64
