"""Graph-replayed step sequences (LLGRunner::run) against the same steps launched one by one
(SPIRIT_HIP_GRAPHS=0): bit-identical spins after every call, with host reads, eager steps and
synchronisations between the replays, for unobserved runs (Simulation_N_Shot), runs observed
every step with the device-side convergence test (Simulation_LLG_Start), and the batched
ensemble."""
import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture
def make_state(tmp_path):
    from spirit_amd import state as st
    made = []

    def make(cfg, k=0):
        path = str(tmp_path / ("in%d.cfg" % len(made)))
        configs.write_config(cfg, path)
        p = st.setup(path, True)
        assert p
        made.append(p)
        return p

    yield make
    for p in made:
        st.delete(p)


def spins(p):
    from spirit_amd import system
    return np.array(system.get_spin_directions(p))


FILM = configs.scaled(configs.FILM_128, [12, 10, 4])  # DDI, Heun, T = 10 K (counter-based noise)


@pytest.mark.parametrize("solver", [3, 2, 4])
def test_n_shot_graph_equals_eager(make_state, monkeypatch, solver):
    from spirit_amd import simulation, system
    print_case("graphs n_shot", solver)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("SPIRIT_HIP_GRAPHS", mode)
        p = make_state(FILM)
        system.get_spin_directions(p)[:] = configs.random_spins(system.get_nos(p), 4)
        simulation.start(p, simulation.METHOD_LLG, solver, single_shot=True)
        seq = []
        for n in (50, 37, 100):  # 3, 2 and 6 replays of 16 steps with eager steps between
            simulation.n_shot(p, n)
            seq.append(spins(p))
            seq.append(system.get_energy(p))
        simulation.stop(p)
        out[mode] = seq
    for a, b in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("amortize", [1, 3, 100])
def test_llg_start_graph_equals_eager(make_state, monkeypatch, amortize):
    """Observed every `amortize` steps (1 and 3: whole periods in the graph; 100: graphs of
    unobserved steps between observations), direct minimisation stopping on convergence."""
    from spirit_amd import simulation, system
    from spirit_amd.parameters import llg
    print_case("graphs llg_start amortize", amortize)
    cfg = dict(configs.scaled(configs.SQUARE_64, [24, 20, 1]), llg_force_convergence=5e-4,
               llg_n_iterations_amortize=amortize)
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("SPIRIT_HIP_GRAPHS", mode)
        p = make_state(cfg)
        system.get_spin_directions(p)[:] = configs.random_spins(system.get_nos(p), 9)
        llg.set_direct_minimization(p, True)
        info = simulation.start(p, simulation.METHOD_LLG, 2, n_iterations=3000)
        res[mode] = (spins(p), system.get_energy(p), info.total_iterations)
    np.testing.assert_array_equal(res["0"][0], res["1"][0])
    assert res["0"][1] == res["1"][1]
    assert res["0"][2] == res["1"][2] < 3000


def test_batch_graph_equals_eager(make_state, monkeypatch):
    from spirit_amd import hip, system
    print_case("graphs batch")
    cfgs = configs.ensemble(6, seed=3)
    for c in cfgs:
        c["n_basis_cells"] = [min(c["n_basis_cells"][0], 48)] * 2 + [1]
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("SPIRIT_HIP_GRAPHS", mode)
        ps = [make_state(c) for c in cfgs]
        for k, p in enumerate(ps):
            system.get_spin_directions(p)[:] = configs.random_spins(system.get_nos(p), 50 + k)
        hip.llg_start_batch(ps, 2, n_iterations=120)
        res[mode] = [spins(p) for p in ps]
    for a, b in zip(res["0"], res["1"]):
        np.testing.assert_array_equal(a, b)
