"""Metropolis Monte Carlo on the GPU (Simulation_MC_Start, Method_MC.cpp:62-185) against
closed forms of classical spins in equilibrium (the random-number streams differ from the
reference's sequential mt19937, so parity is statistical):
* a paramagnet in a field: <s_z> = L(x), x = mu mu_B B / (k_B T), L(x) = coth x - 1/x;
* the classical Heisenberg chain (Fisher 1964): <s_i . s_i+1> = L(J / (k_B T)).
Tolerances are five standard errors of the one-configuration estimate (N spins or bonds,
variance below 1/3 per term), and the T = 0 sweep never raises the energy."""
import math

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

K_B = 0.08617330350
MU_B = 0.057883817555


def print_case(*a):
    print("CASE", *a, flush=True)


def langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


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


def run_mc(p, T, n, cone=30.0, adaptive=True):
    from spirit_amd import simulation
    from spirit_amd.parameters import mc
    mc.set_temperature(p, T)
    mc.set_metropolis_cone(p, True, cone, adaptive, 0.5)
    info = simulation.start(p, simulation.METHOD_MC, n_iterations=n)
    assert info.total_iterations == n


def test_mc_paramagnet_langevin(make_state):
    print_case("mc paramagnet")
    N = 256 * 256
    cfg = dict(bravais_lattice="sc", n_basis_cells=[256, 256, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=1.0, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.0,
               n_shells_exchange=0, n_shells_dmi=0)
    p = make_state(cfg)
    from spirit_amd import system
    for x in (0.5, 2.0):
        T = 2.0 * MU_B * 1.0 / (K_B * x)
        run_mc(p, T, 400)
        s = np.array(system.get_spin_directions(p))
        np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-12)
        mz = s[:, 2].mean()
        assert abs(mz - langevin(x)) < 5 * math.sqrt(1.0 / 3.0 / N), (x, mz, langevin(x))


def test_mc_heisenberg_chain(make_state):
    print_case("mc chain")
    N = 65536
    cfg = dict(bravais_lattice="sc", n_basis_cells=[N, 1, 1], boundary_conditions=[1, 0, 0], mu_s=[2.0],
               external_field_magnitude=0.0, anisotropy_magnitude=0.0, n_shells_exchange=1, jij=[1.0],
               n_shells_dmi=0)
    p = make_state(cfg)
    from spirit_amd import system
    for K in (0.5, 2.0):
        T = 1.0 / (K_B * K)
        run_mc(p, T, 1500)
        s = np.array(system.get_spin_directions(p))
        c = np.einsum("ij,ij->i", s, np.roll(s, -1, axis=0)).mean()
        assert abs(c - langevin(K)) < 5 * math.sqrt(1.0 / 3.0 / N) * 3, (K, c, langevin(K))


def test_mc_zero_temperature_and_cone(make_state, oracle):
    """T = 0: every accepted trial lowers the single-spin energy, so the total energy of the
    Neel-DMI square lattice never rises; the adaptive cone moves away from its start."""
    from spirit_amd import system
    from spirit_amd.parameters import mc
    print_case("mc T=0")
    cfg = configs.scaled(configs.SQUARE_64, [32, 32, 1])
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 5)
    system.get_spin_directions(p)[:] = s0
    _, E_prev = S.gradient_and_energy(s0)
    for _ in range(5):
        run_mc(p, 0.0, 20)
        s = np.array(system.get_spin_directions(p))
        _, E = S.gradient_and_energy(s)
        assert E <= E_prev + 1e-9
        E_prev = E
    assert abs(system.get_energy(p) - E_prev) <= 1e-5 * abs(E_prev) + 1e-3
    cone = mc.get_metropolis_cone(p)
    assert cone[0] and cone[2] and abs(cone[1] - 30.0) > 0.5


def test_mc_single_shot_equals_full_run(make_state):
    """Simulation_MC_Start in single-shot mode, then Simulation_N_Shot / _SingleShot (the
    reference's Simulation.cpp:454-535): 25 + 20 + 15 single iterations give the spins, the
    adapted cone angle and the iteration count of one 60-iteration run, bit for bit (same
    kernels, same counter-based random numbers); the method ends by itself after the 60th
    iteration, and Simulation_Stop ends a single-shot run early."""
    from spirit_amd import simulation, system
    from spirit_amd.parameters import mc
    print_case("mc single shot")
    cfg = dict(bravais_lattice="sc", n_basis_cells=[40, 30, 2], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=0.5, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.1,
               anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[3.0], n_shells_dmi=1, dij=[1.0],
               dm_chirality=2)
    runs = []
    for single in (False, True):
        p = make_state(cfg)
        s0 = configs.random_spins(system.get_nos(p), 71)
        system.get_spin_directions(p)[:] = s0
        mc.set_temperature(p, 15.0)
        mc.set_metropolis_cone(p, True, 25.0, True, 0.5)
        if single:
            simulation.start(p, simulation.METHOD_MC, n_iterations=60, single_shot=True)
            assert simulation.running_on_image(p)
            assert simulation.get_method_name(p) == "MC"
            simulation.n_shot(p, 25)
            simulation.n_shot(p, 20)
            assert simulation.get_iteration(p) == 45
            mid = np.array(system.get_spin_directions(p))
            assert np.abs(mid - s0).max() > 1e-3  # the host spins follow the run
            for _ in range(15):
                simulation.single_shot(p)
            assert not simulation.running_on_image(p)  # ended after the 60th iteration
        else:
            info = simulation.start(p, simulation.METHOD_MC, n_iterations=60)
            assert info.total_iterations == 60
        runs.append((np.array(system.get_spin_directions(p)), mc.get_metropolis_cone(p)))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]
    # Simulation_Stop ends a single-shot run; the spins stay those of the iterations done
    p = make_state(cfg)
    system.get_spin_directions(p)[:] = configs.random_spins(system.get_nos(p), 72)
    simulation.start(p, simulation.METHOD_MC, n_iterations=100, single_shot=True)
    simulation.n_shot(p, 10)
    after10 = np.array(system.get_spin_directions(p))
    simulation.stop(p)
    assert not simulation.running_on_image(p)
    np.testing.assert_array_equal(np.array(system.get_spin_directions(p)), after10)
