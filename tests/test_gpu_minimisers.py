"""Direct-minimisation solvers of the LLG method on the GPU (VP, VP_OSO, LBFGS_OSO,
LBFGS_Atlas) through the C API: trajectories against the oracle (1e-9 per component after
N iterations; the reductions run in a different order, hence not bit-exact), the reference's
minimum (core/test/test_solvers.cpp:40-72) for all eight LLG solvers, and graph replays
bit-identical to launches one by one."""
import json
import os

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))

MINIMISERS = {"vp": 0, "lbfgs_oso": 5, "lbfgs_atlas": 6, "vp_oso": 7}
ALL_SOLVERS = {"vp": 0, "sib": 1, "depondt": 2, "heun": 3, "rk4": 4, "lbfgs_oso": 5, "lbfgs_atlas": 6,
               "vp_oso": 7}


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture
def make_state(tmp_path):
    from spirit_amd import state
    made = []

    def make(cfg):
        path = str(tmp_path / ("input%d.cfg" % len(made)))
        configs.write_config(cfg, path)
        p = state.setup(path, True)
        assert p
        made.append(p)
        return p

    yield make
    for p in made:
        state.delete(p)
    os.environ.pop("SPIRIT_HIP_GRAPHS", None)


def run_n(p, s0, solver, n):
    from spirit_amd import simulation, system
    from spirit_amd.parameters import llg
    llg.set_direct_minimization(p, True)
    system.get_spin_directions(p)[:] = s0
    simulation.start(p, simulation.METHOD_LLG, solver, single_shot=True)
    simulation.n_shot(p, n)
    s = np.array(system.get_spin_directions(p))
    E = system.get_energy(p)
    T = simulation.get_max_torque_norm(p)
    simulation.stop(p)
    return s, E, T


CASES = {
    # the solvers test lattice (16x16 periodic square, J, Bloch DMI, B = 25 T)
    "square16": (lambda: dict(GOLD["solvers"]["cfg"]), 60),
    # open film with DDI (the DDI field through the FFT passes every iteration)
    "film_ddi": (lambda: configs.scaled(configs.FILM_128, [12, 10, 4]), 40),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("name", list(MINIMISERS))
def test_minimiser_trajectory(make_state, oracle, name, case):
    cfg_fn, n = CASES[case]
    cfg = cfg_fn()
    print_case("minimiser", name, case)
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 41)
    s, E, T = run_n(p, s0, MINIMISERS[name], n)
    s_ref, E_ref, T_ref, _ = S.llg(s0, MINIMISERS[name], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"],
                                   direct_minimization=True)
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()
    assert abs(E - E_ref) <= 1e-5 * abs(E_ref) + 1e-3  # System_Get_Energy returns a float
    assert abs(T - T_ref) <= 1e-6 * T_ref + 1e-9


@pytest.mark.parametrize("name", list(ALL_SOLVERS))
def test_all_llg_solvers_reach_reference_minimum(make_state, name):
    """core/test/test_solvers.cpp:40-72: Simulation_LLG_Start with direct minimisation from the
    skyrmion, every LLG solver of the reference."""
    from spirit_amd import configuration, quantities, simulation, system
    from spirit_amd.parameters import llg
    G = GOLD["solvers"]
    print_case("solvers golden", name)
    p = make_state(G["cfg"])
    llg.set_direct_minimization(p, True)
    configuration.plus_z(p)
    sk = G["skyrmion"]
    configuration.skyrmion(p, sk["r"], sk["order"], sk["phase"], sk["upDown"], sk["achiral"], sk["rl"])
    simulation.start(p, simulation.METHOD_LLG, ALL_SOLVERS[name])
    E = system.get_energy(p)
    m = quantities.get_magnetization(p)
    assert abs(E - G["energy_minimum"]) < G["tolerance"], (name, E)
    assert abs(m[2] - G["mz_minimum"]) < G["tolerance"], (name, m)
    assert simulation.get_max_torque_norm(p) < G["cfg"]["llg_force_convergence"]


@pytest.mark.parametrize("name", ["lbfgs_atlas", "vp"])
def test_minimiser_graph_replay_equals_eager(make_state, oracle, name):
    """A run of 100 iterations replays captured graphs of 16; with SPIRIT_HIP_GRAPHS=0 every
    kernel is launched one by one.  The spins agree bit for bit."""
    cfg = dict(GOLD["solvers"]["cfg"])
    print_case("minimiser graphs", name)
    p = make_state(cfg)
    s0 = oracle.random_spins(16 * 16, 43)
    os.environ["SPIRIT_HIP_GRAPHS"] = "0"
    s_eager, _, _ = run_n(p, s0, MINIMISERS[name], 100)
    os.environ["SPIRIT_HIP_GRAPHS"] = "1"
    s_graph, _, _ = run_n(p, s0, MINIMISERS[name], 100)
    assert np.array_equal(s_graph, s_eager), np.abs(s_graph - s_eager).max()
