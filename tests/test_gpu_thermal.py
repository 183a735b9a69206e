"""The engine's stochastic LLG against equilibrium statistics (GPU): free spins in a field
B at temperature T, advanced by Heun with the engine's counter-based thermal field, reach
the Boltzmann distribution, <m_z> = L(mu_s mu_B B / k_B T), L(x) = coth x - 1/x.  This fixes
the thermal amplitude eps = sqrt(2 alpha dt gamma k_B / mu_B) / (1 + alpha^2) sqrt(T / mu_s)
of Method_LLG::Prepare_Thermal_Field (Method_LLG.cpp:66-110) on the GPU path, as the oracle's
test does on the CPU: a factor 2 in the noise variance moves <m_z> by 46%."""
import math

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

K_B = 0.08617330350
MU_B = 0.057883817555


def langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


@pytest.mark.parametrize("temperature", [20.0, 60.0])
def test_llg_paramagnet_langevin(tmp_path, temperature):
    from spirit_amd import simulation, state, system
    print("CASE llg paramagnet T=%g" % temperature, flush=True)
    cfg = dict(bravais_lattice="sc", n_basis_cells=[128, 128, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=20.0, external_field_normal=[0, 0, 1], n_shells_exchange=0,
               n_shells_dmi=0, ddi_method="none", llg_dt=0.004, llg_damping=0.3, llg_seed=31,
               llg_temperature=temperature)
    path = str(tmp_path / "para.cfg")
    configs.write_config(cfg, path)
    p = state.setup(path, True)
    assert p
    try:
        nos = system.get_nos(p)
        # random start: spins at rest along the field have zero torque, and a method that starts
        # converged takes no step (Method::ContinueIterating)
        system.get_spin_directions(p)[:] = configs.random_spins(nos, 5)
        simulation.start(p, simulation.METHOD_LLG, simulation.SOLVER_HEUN, single_shot=True)
        simulation.n_shot(p, 3000)  # 12 ps: about ten relaxation times
        m = []
        for _ in range(40):
            simulation.n_shot(p, 750)  # samples 3 ps apart
            s = np.array(system.get_spin_directions(p))
            m.append(float(s[:, 2].mean()))
        simulation.stop(p)
    finally:
        state.delete(p)
    m = np.array(m)
    mean, se = float(m.mean()), float(m.std(ddof=1) / math.sqrt(len(m)))
    want = langevin(2.0 * MU_B * 20.0 / (K_B * temperature))
    print("  <m_z> = %.5f +- %.5f, Langevin %.5f" % (mean, se, want), flush=True)
    assert abs(mean - want) < max(5 * se, 0.01), (mean, want, se)
