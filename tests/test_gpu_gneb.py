"""GPU parity of the GNEB method: the HIP engine (through the C API) against the oracle's
restatement of Method_GNEB on the same chain, and the reference's saddle-point known answer
(core/test/test_solvers.cpp:74-133) reached through the API.

Tolerances: chain spins after N iterations 1e-9 max-abs per component; energies to float
rounding (Chain_Get_Energy returns floats); reaction coordinates 1e-6 relative.  The GPU forms dots with normalised tangents as (1/|tau|)(F . tau_raw) and
sums in a different order than the oracle, so agreement is to rounding, not bit-exact."""
import json
import os

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))
SOLVERS = {"vp": 0, "sib": 1, "depondt": 2, "heun": 3, "rk4": 4, "lbfgs_oso": 5, "lbfgs_atlas": 6, "vp_oso": 7}


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture
def state(tmp_path):
    from spirit_amd import state as st
    made = []

    def make(cfg):
        path = str(tmp_path / "input.cfg")
        configs.write_config(cfg, path)
        p = st.setup(path, True)
        assert p
        made.append(p)
        return p

    yield make
    for p in made:
        st.delete(p)


def collapse_chain(p, noi, sk):
    """Skyrmion on image 0, +z on the last image, homogeneous rotation in between (all via the API)."""
    from spirit_amd import chain, configuration, transition
    configuration.plus_z(p)
    configuration.skyrmion(p, sk["r"], sk["order"], sk["phase"], sk["upDown"], sk["achiral"], sk["rl"])
    chain.image_to_clipboard(p)
    for _ in range(noi - 1):
        chain.insert_image_after(p)
    chain.jump_to_image(p, noi - 1)
    configuration.plus_z(p)
    chain.jump_to_image(p, 0)
    transition.homogeneous(p, 0, noi - 1)


def chain_spins(p, noi):
    from spirit_amd import system
    return np.array([np.array(system.get_spin_directions(p, idx_image=k)) for k in range(noi)])


SMALL = dict(GOLD["solvers"]["cfg"], n_basis_cells=[12, 12, 1], gneb_force_convergence=1e-14,
             external_field_magnitude=20.0)
SK = dict(r=4.0, order=1.0, phase=-90.0, upDown=False, achiral=False, rl=False)


def run_both(state, oracle, cfg, solver, n_iter, noi=7, types=None, setup=None, oracle_kw=None, perturb=0.0):
    from spirit_amd import chain, simulation, system
    from spirit_amd.parameters import gneb
    p = state(cfg)
    collapse_chain(p, noi, SK)
    c0 = chain_spins(p, noi)
    S = oracle.System(cfg)
    # the engine's homogeneous rotation is the oracle's (Configuration_Chain.cpp:28-75)
    ref0 = np.zeros_like(c0)
    ref0[0], ref0[-1] = c0[0], c0[-1]
    ref0 = oracle.homogeneous_rotation(ref0, 0, noi - 1)
    assert np.abs(ref0 - c0).max() < 1e-14
    if perturb:
        # seeded noise on the interior images (written through the spin mirror)
        rng = np.random.default_rng(3)
        for k in range(1, noi - 1):
            v = c0[k] + perturb * rng.normal(size=c0[k].shape)
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            system.get_spin_directions(p, idx_image=k)[:] = v
            c0[k] = v
    if types is not None:
        for k, t in enumerate(types):
            gneb.set_climbing_falling(p, int(t), idx_image=k)
    if setup:
        setup(p)
    simulation.start(p, simulation.METHOD_GNEB, solver, n_iterations=n_iter)
    c_gpu = chain_spins(p, noi)
    E_gpu = np.array(chain.get_energy(p))
    Rx_gpu = np.array(chain.get_rx(p))
    kw = dict(dt=cfg["llg_dt"], spring_constant=cfg.get("gneb_spring_constant", 1.0),
              force_convergence=cfg["gneb_force_convergence"])
    kw.update(oracle_kw or {})
    c_ref, data, n = S.gneb(c0, solver, n_iter, image_type=types, **kw)
    assert n == n_iter
    return p, c_gpu, c_ref, E_gpu, Rx_gpu, data


def check(c_gpu, c_ref, E_gpu, Rx_gpu, data, tol=1e-9):
    err = np.abs(c_gpu - c_ref).max()
    assert err < tol, err
    np.testing.assert_allclose(E_gpu, data["energies"].astype(np.float32), rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(Rx_gpu, data["Rx"], rtol=1e-6, atol=1e-6)


# iterations compared per solver: where a solver amplifies rounding differences the run is
# shortened (measured on the oracle alone with 1e-15 perturbations of the interior images:
# VP 1e-12 after 20 and 4e-10 after 100 iterations; VP_OSO 4e-13 after 60 and 4e-8 after 100;
# LBFGS_Atlas 3e-13 after 60; LBFGS_OSO 2e-14 after 100)
TRAJECTORY_ITERATIONS = {"vp": 60, "vp_oso": 40, "lbfgs_atlas": 60, "lbfgs_oso": 100}


@pytest.mark.parametrize("name", ["vp", "heun", "depondt", "sib", "rk4", "lbfgs_oso", "lbfgs_atlas", "vp_oso"])
def test_gneb_trajectory(state, oracle, name):
    """Every solver on a 7-image skyrmion-collapse chain with a climbing, a falling and a
    stationary image.  The predictor-corrector solvers agree to ~1e-15 after 300 iterations;
    the velocity-projection and L-BFGS solvers over the iterations of TRAJECTORY_ITERATIONS."""
    print_case("gneb trajectory", name)
    n = TRAJECTORY_ITERATIONS.get(name, 300)
    types = [0, 0, 1, 0, 2, 3, 0]
    _, cg, cr, Eg, Rg, data = run_both(state, oracle, SMALL, SOLVERS[name], n, types=types)
    check(cg, cr, Eg, Rg, data)


def test_gneb_energy_weighted_springs(state, oracle):
    """spring_force_ratio > 0: springs from the Hermite-interpolated path length (Heun)."""
    from spirit_amd.parameters import gneb
    print_case("gneb", "spring_force_ratio 0.4")
    _, cg, cr, Eg, Rg, data = run_both(
        state, oracle, SMALL, 3, 200, setup=lambda p: gneb.set_spring_force(p, 1.0, 0.4),
        oracle_kw=dict(spring_force_ratio=np.float32(0.4)))
    check(cg, cr, Eg, Rg, data)


def test_gneb_path_shortening(state, oracle):
    """The shrinking force normalises the projected difference of the unit image differences
    (Method_GNEB.cpp:204-228); on the evenly spaced geodesic chain that difference is rounding
    noise, so a 1e-16 change flips its direction (1e-8 after 200 iterations, in the oracle
    alone).  Interior images start perturbed by seeded noise (amplitude 0.05), where the
    same 1e-16 change stays below 1e-12."""
    print_case("gneb", "path shortening 0.05")
    from spirit_amd.parameters import gneb
    _, cg, cr, Eg, Rg, data = run_both(
        state, oracle, SMALL, 3, 200, setup=lambda p: gneb.set_path_shortening_constant(p, 0.05),
        oracle_kw=dict(path_shortening_constant=np.float32(0.05)), perturb=0.05)
    check(cg, cr, Eg, Rg, data)


@pytest.mark.parametrize("translating,escape", [(False, False), (True, False), (True, True)])
def test_gneb_moving_endpoints(state, oracle, translating, escape):
    """Moving endpoints (set through the API), translating endpoints, and escape_first (set
    through the input file: the reference has no setter for it)."""
    print_case("gneb moving endpoints", translating, escape)
    from spirit_amd.parameters import gneb

    def setup(p):
        gneb.set_moving_endpoints(p, True)
        gneb.set_translating_endpoints(p, translating)
        gneb.set_equilibrium_delta_Rx(p, 0.8, 1.2)

    cfg = dict(SMALL, gneb_escape_first=int(escape))
    _, cg, cr, Eg, Rg, data = run_both(
        state, oracle, cfg, 3, 150, setup=setup,
        oracle_kw=dict(moving_endpoints=True, translating_endpoints=translating,
                       delta_Rx=(float(np.float32(0.8)), float(np.float32(1.2))), escape_first=escape))
    check(cg, cr, Eg, Rg, data)


def test_gneb_with_ddi(state, oracle):
    """A chain whose images carry the dipolar FFT field (open 8x8x2 film)."""
    print_case("gneb", "ddi film")
    cfg = dict(SMALL, n_basis_cells=[8, 8, 2], boundary_conditions=[0, 0, 0], ddi_method="fft",
               ddi_n_periodic_images=[0, 0, 0], ddi_pb_zero_padding=1)
    _, cg, cr, Eg, Rg, data = run_both(state, oracle, cfg, 3, 100, noi=5)
    check(cg, cr, Eg, Rg, data)


def test_gneb_single_shot_matches_full_run(state):
    """Simulation_N_Shot on a single-shot GNEB advances the same chain as a blocking start."""
    from spirit_amd import simulation
    print_case("gneb", "single shot")
    p1 = state(SMALL)
    collapse_chain(p1, 5, SK)
    simulation.start(p1, simulation.METHOD_GNEB, 0, n_iterations=40)
    a = chain_spins(p1, 5)
    p2 = state(SMALL)
    collapse_chain(p2, 5, SK)
    simulation.start(p2, simulation.METHOD_GNEB, 0, n_iterations=40, single_shot=True)
    for _ in range(4):
        simulation.n_shot(p2, 10)
    b = chain_spins(p2, 5)
    simulation.stop(p2)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", ["lbfgs_atlas", "lbfgs_oso", "vp_oso", "vp", "heun", "depondt"])
def test_gneb_saddle_point_golden_api(state, name):
    """core/test/test_solvers.cpp:74-133 through the API: LLG minimisation of a skyrmion,
    9-image chain, 2e4 GNEB iterations, climbing image set automatically, GNEB to convergence;
    the highest inner image has the reference's energy and magnetisation (float, 1e-5)."""
    from spirit_amd import chain, configuration, quantities, simulation, system, transition
    from spirit_amd.parameters import gneb, llg
    print_case("gneb saddle golden", name)
    G = GOLD["solvers"]
    p = state(G["cfg"])
    llg.set_direct_minimization(p, True)
    configuration.plus_z(p)
    sk = G["skyrmion"]
    configuration.skyrmion(p, sk["r"], sk["order"], sk["phase"], sk["upDown"], sk["achiral"], sk["rl"])
    simulation.start(p, simulation.METHOD_LLG, 4)
    assert abs(system.get_energy(p) - G["energy_minimum"]) < G["tolerance"]
    noi = G["n_images"]
    chain.image_to_clipboard(p)
    for _ in range(1, noi):
        chain.insert_image_after(p)
    chain.replace_image(p, 0)
    chain.jump_to_image(p, noi - 1)
    configuration.plus_z(p)
    chain.jump_to_image(p, 0)
    transition.homogeneous(p, 0, noi - 1)
    simulation.start(p, simulation.METHOD_GNEB, SOLVERS[name], n_iterations=20000)
    gneb.set_image_type_automatically(p)
    simulation.start(p, simulation.METHOD_GNEB, SOLVERS[name])
    i_max, E_max = 1, system.get_energy(p, idx_image=0)
    for i in range(1, noi - 1):
        Ei = system.get_energy(p, idx_image=i)
        if Ei > E_max:
            i_max, E_max = i, Ei
    m = quantities.get_magnetization(p, idx_image=i_max)
    assert abs(E_max - G["energy_saddle"]) < G["tolerance"], (name, E_max)
    assert abs(m[2] - G["mz_saddle"]) < G["tolerance"], (name, m)
    assert gneb.get_climbing_falling(p, idx_image=i_max) == gneb.IMAGE_CLIMBING
    assert max(simulation.get_chain_max_torque_norms(p, noi)) < G["cfg"]["gneb_force_convergence"]


def test_gneb_hexagonal_climbing_vp(state, oracle):
    """The GNEB workload's lattice and Hamiltonian at a reduced size: a 24x24 hexagonal
    monolayer (J=10, Neel DMI 6, K=0.4, B=3 T), 9 images from a Neel skyrmion to the
    ferromagnet, climbing image in the middle, VP, 40 iterations."""
    print_case("gneb", "hex 24x24 climbing vp")
    cfg = dict(configs.HEX_128, n_basis_cells=[24, 24, 1], gneb_force_convergence=1e-14)
    noi = 9
    types = [0] * noi
    types[4] = 1
    _, cg, cr, Eg, Rg, data = run_both(state, oracle, cfg, SOLVERS["vp"], 40, noi=noi, types=types)
    check(cg, cr, Eg, Rg, data)


def test_gneb_full_size_hexagonal_chain(state, oracle):
    """The GNEB workload at its full size: a 128x128 hexagonal monolayer (J=10, Neel DMI 6,
    K=0.4, B=3 T), 32 images from a Neel skyrmion to the ferromagnet, climbing image in the
    middle, VP, 20 iterations against the oracle (VP amplifies rounding differences to about
    1e-12 over 20 iterations)."""
    print_case("gneb", "hex 128x128 32 images climbing vp")
    cfg = dict(configs.HEX_128, gneb_force_convergence=1e-14)
    noi = 32
    types = [0] * noi
    types[16] = 1
    _, cg, cr, Eg, Rg, data = run_both(state, oracle, cfg, SOLVERS["vp"], 20, noi=noi, types=types)
    check(cg, cr, Eg, Rg, data)
