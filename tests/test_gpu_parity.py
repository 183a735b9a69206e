"""GPU parity: the HIP engine (through the C API) against the CPU oracle on the same
seeded inputs.  Tolerances: gradient 1e-12 relative, energy 1e-11 relative, spins
after N steps 1e-9 max-abs per component (FP64 throughout; the GPU contracts
multiply-adds into FMAs, the oracle does not, hence not bit-exact)."""
import json
import math
import os

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))

SOLVERS = {"depondt": 2, "heun": 3, "sib": 1, "rk4": 4}


class Engine:
    """A State built from a config dict, with helpers to read/write its spins."""

    def __init__(self, cfg, tmp_path, site_K=None):
        from spirit_amd import state
        path = str(tmp_path / "input.cfg")
        configs.write_config(cfg, path)
        self.p = state.setup(path, True)
        assert self.p
        from spirit_amd import system
        self.nos = system.get_nos(self.p)
        if site_K is not None:
            from spirit_amd import hip
            hip.set_site_anisotropy(self.p, site_K, cfg.get("site_anisotropy_normal", (0, 0, 1)))

    def set_spins(self, s):
        from spirit_amd import system
        v = system.get_spin_directions(self.p)
        v[:] = s

    def spins(self):
        from spirit_amd import system
        return np.array(system.get_spin_directions(self.p))

    def gradient(self):
        from spirit_amd import hip
        return hip.gradient_and_energy(self.p, self.nos)

    def close(self):
        from spirit_amd import state
        state.delete(self.p)


@pytest.fixture
def engine(tmp_path):
    made = []

    def make(cfg, site_K=None):
        e = Engine(cfg, tmp_path, site_K)
        made.append(e)
        return e

    yield make
    for e in made:
        e.close()


def print_case(*a):
    print("CASE", *a, flush=True)


GRAD_CASES = {
    "square_neel_periodic": configs.scaled(configs.SQUARE_64, [32, 24, 1]),
    "film_bloch_ddi_open": configs.scaled(configs.FILM_128, [12, 10, 4]),
    "hex_neel_periodic": dict(bravais_lattice="hex2d", n_basis_cells=[20, 16, 1], boundary_conditions=[1, 1, 0],
                              mu_s=[2.0], external_field_magnitude=3.0, anisotropy_magnitude=0.4, n_shells_exchange=1,
                              jij=[10.0], n_shells_dmi=1, dij=[6.0], dm_chirality=2),
    "skewed_two_basis_ddi_periodic_images": GOLD["physics_ddi"]["cfg"],
    "pairs_open": GOLD["fd_pairs"]["cfg"],
    "cubic_two_shells_mixed_bc": dict(bravais_lattice="sc", n_basis_cells=[9, 7, 5], boundary_conditions=[1, 0, 1],
                                      mu_s=[1.5], external_field_magnitude=1.0, external_field_normal=[1, 2, 3],
                                      anisotropy_magnitude=0.3, anisotropy_normal=[1, 0, 1],
                                      cubic_anisotropy_magnitude=0.1, n_shells_exchange=2, jij=[3.0, -1.0],
                                      n_shells_dmi=1, dij=[0.7], dm_chirality=-1),
    "ddi_periodic_no_padding": dict(bravais_lattice="sc", n_basis_cells=[8, 6, 3], boundary_conditions=[1, 1, 0],
                                    mu_s=[2.0], n_shells_exchange=1, jij=[1.0], ddi_method="fft",
                                    ddi_n_periodic_images=[2, 1, 0], ddi_pb_zero_padding=0),
    "ddi_odd_sizes": dict(bravais_lattice="sc", n_basis_cells=[7, 5, 3], boundary_conditions=[0, 0, 0], mu_s=[2.0],
                          ddi_method="fft", ddi_n_periodic_images=[0, 0, 0]),
    # real-space methods (ddi_method cutoff): pairs within the radius wrapping periodic axes,
    # and the direct sum (negative radius) with periodic images
    "ddi_cutoff_periodic": dict(bravais_lattice="sc", n_basis_cells=[9, 8, 3], boundary_conditions=[1, 1, 0],
                                mu_s=[2.0], external_field_magnitude=0.5, n_shells_exchange=1, jij=[2.0],
                                n_shells_dmi=1, dij=[0.5], dm_chirality=1, ddi_method="cutoff", ddi_radius=2.5),
    "ddi_cutoff_two_basis_open": dict(GOLD["physics_ddi"]["cfg"], boundary_conditions=[0, 0, 0], ddi_method="cutoff",
                                      ddi_radius=3.1),
    "ddi_direct_periodic_images": dict(bravais_lattice="sc", n_basis_cells=[6, 5, 2], boundary_conditions=[1, 1, 0],
                                       mu_s=[2.0], n_shells_exchange=1, jij=[1.0], ddi_method="cutoff",
                                       ddi_radius=-1.0, ddi_n_periodic_images=[2, 1, 3]),
    "ddi_direct_hex_open": dict(bravais_lattice="hex2d", n_basis_cells=[17, 13, 1], boundary_conditions=[0, 0, 0],
                                mu_s=[1.5], anisotropy_magnitude=0.2, ddi_method="cutoff", ddi_radius=-1.0),
    # four-spin interactions (n_interaction_quadruplets) on a two-atom basis, periodic and open axes
    "quadruplets_two_basis": dict(bravais_lattice="sc", n_basis_cells=[6, 5, 4], boundary_conditions=[1, 0, 1],
                                  mu_s=[1.0, 2.0], basis=[[0, 0, 0], [0.5, 0.5, 0.5]], n_shells_exchange=1,
                                  jij=[2.0], external_field_magnitude=0.3,
                                  quadruplets=[[0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 3.0],
                                               [0, 1, 0, 0, 0, 1, 1, 0, 0, 0, -1, 0, 0, -1.25],
                                               [1, 1, 0, 0, -1, 0, 2, 0, 0, 1, 0, -1, 1, 0.5]]),
}


@pytest.mark.parametrize("name", list(GRAD_CASES.keys()))
def test_gradient_and_energy(engine, oracle, name):
    cfg = GRAD_CASES[name]
    print_case("gradient", name)
    e = engine(cfg)
    S = oracle.System(cfg)
    assert e.nos == S.nos
    s = oracle.random_spins(S.nos, 1234)
    e.set_spins(s)
    g, E = e.gradient()
    g0, E0 = S.gradient_and_energy(s)
    scale = np.abs(g0).max()
    assert np.abs(g - g0).max() <= 1e-12 * scale, np.abs(g - g0).max() / scale
    assert abs(E - E0) <= 1e-11 * max(1.0, abs(E0))


@pytest.mark.parametrize("kind,code", [("generic", 0), ("lds", 1), ("reg", 2), ("reg4", 3)])
def test_ddi_pass_c_kernels(engine, oracle, monkeypatch, kind, code):
    """Every pass C kernel (forced through SPIRIT_HIP_DDI_PASS_C) gives the oracle's dipolar
    gradient on a 16x12x128 film (padded 32x24x256: power-of-two c columns, mixed-radix b)."""
    from spirit_amd import hip, simulation
    print_case("ddi pass C kernel", kind)
    monkeypatch.setenv("SPIRIT_HIP_DDI_PASS_C", kind)
    cfg = configs.scaled(configs.FILM_128, [16, 12, 128])
    e = engine(cfg)
    S = oracle.System(cfg)
    s = oracle.random_spins(S.nos, 4321)
    e.set_spins(s)
    simulation.start(e.p, simulation.METHOD_LLG, 2, single_shot=True)
    info = hip.ddi_info(e.p)
    simulation.stop(e.p)
    assert ("pass C kernel %d " % code) in info, info
    g, E = e.gradient()
    g0, E0 = S.gradient_and_energy(s)
    assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    assert abs(E - E0) <= 1e-11 * abs(E0)


def test_gradient_site_anisotropy(engine, oracle):
    cfg = configs.scaled(configs.CUBE_256, [16, 12, 6])
    print_case("gradient", "site anisotropy + ddi")
    K = configs.site_anisotropy(16 * 12 * 6, 7)
    e = engine(cfg, site_K=K)
    S = oracle.System(cfg, site_K=K)
    s = oracle.random_spins(S.nos, 99)
    e.set_spins(s)
    g, E = e.gradient()
    g0, E0 = S.gradient_and_energy(s)
    assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    assert abs(E - E0) <= 1e-11 * abs(E0)


def _run_llg(e, solver, n, chunk=None):
    from spirit_amd import simulation
    simulation.start(e.p, simulation.METHOD_LLG, solver, single_shot=True)
    simulation.n_shot(e.p, n)
    s = e.spins()
    from spirit_amd import system
    E = system.get_energy(e.p)
    T = simulation.get_max_torque_norm(e.p)
    simulation.stop(e.p)
    return s, E, T


@pytest.mark.parametrize("solver", list(SOLVERS.keys()))
def test_llg_trajectory_square(engine, oracle, solver):
    cfg = configs.scaled(configs.SQUARE_64, [32, 32, 1])
    print_case("llg", solver, "32x32 neel")
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 20006)
    e.set_spins(s0)
    n = 300
    s, E, T = _run_llg(e, SOLVERS[solver], n)
    s_ref, E_ref, T_ref, _ = S.llg(s0, SOLVERS[solver], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()
    assert abs(E - E_ref) <= 1e-5 * abs(E_ref) + 1e-3  # System_Get_Energy returns a float
    assert abs(T - T_ref) <= 1e-5 * T_ref + 1e-7


def test_llg_depondt_64x64_2000_steps(engine, oracle):
    cfg = configs.SQUARE_64
    print_case("llg", "depondt", "64x64 2000 steps")
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 20006)
    e.set_spins(s0)
    s, _, _ = _run_llg(e, 2, 2000)
    s_ref, E_ref, _, _ = S.llg(s0, 2, 2000, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()
    # total energy of the final configuration, FP64 on both sides
    e.set_spins(s)
    _, E = e.gradient()
    _, E0 = S.gradient_and_energy(s_ref)
    assert abs(E - E0) <= 1e-11 * abs(E0)


@pytest.mark.parametrize("solver", ["heun", "depondt"])
def test_llg_ddi_trajectory(engine, oracle, solver):
    cfg = configs.scaled(configs.FILM_128, [16, 12, 6], llg_temperature=0.0)
    print_case("llg ddi", solver)
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 5)
    e.set_spins(s0)
    n = 60
    s, _, _ = _run_llg(e, SOLVERS[solver], n)
    s_ref, _, _, _ = S.llg(s0, SOLVERS[solver], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()


@pytest.mark.parametrize("radius", [2.2, -1.0])
def test_llg_ddi_real_space_trajectory(engine, oracle, radius):
    """Heun with the real-space dipolar field (cutoff pairs within 2.2, and the direct sum) on
    a 10x8x3 open film: spins after 40 steps against the oracle."""
    cfg = dict(configs.scaled(configs.FILM_128, [10, 8, 3], llg_temperature=0.0), ddi_method="cutoff",
               ddi_radius=radius)
    print_case("llg ddi real space", radius)
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 15)
    e.set_spins(s0)
    s, _, _ = _run_llg(e, SOLVERS["heun"], 40)
    s_ref, _, _, _ = S.llg(s0, SOLVERS["heun"], 40, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()


def test_llg_quadruplets_trajectory(engine, oracle):
    """Depondt with four-spin interactions, exchange and a field: spins after 100 steps, and a
    lattice with quadruplets only, whose Gradient_and_Energy leaves them out
    (Hamiltonian_Heisenberg.cpp:754), so the spins only precess about nothing: they stay put."""
    cfg = dict(GRAD_CASES["quadruplets_two_basis"], llg_dt=1e-3, llg_damping=0.3, llg_temperature=0.0)
    print_case("llg quadruplets")
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 8)
    e.set_spins(s0)
    s, _, _ = _run_llg(e, SOLVERS["depondt"], 100)
    s_ref, _, _, _ = S.llg(s0, SOLVERS["depondt"], 100, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()
    assert np.abs(s - s0).max() > 1e-3
    only = dict(bravais_lattice="sc", n_basis_cells=[5, 4, 3], boundary_conditions=[1, 1, 1], mu_s=[1.0],
                external_field_magnitude=0.0, quadruplets=[cfg["quadruplets"][0]], llg_dt=1e-3, llg_damping=0.3)
    e2 = engine(only)
    s1 = oracle.random_spins(e2.nos, 9)
    e2.set_spins(s1)
    s, _, _ = _run_llg(e2, SOLVERS["depondt"], 20)
    np.testing.assert_array_equal(s, s1)


def test_llg_ddi_trajectory_row_tiles(engine, oracle):
    """A 256x16x4 film with per-site K_i: rows of 256 cells fill whole blocks, so the stage
    kernels visit the lattice in 16-row tiles swept through c, and most sites take their
    neighbours as fixed index offsets; Depondt + DDI, 30 steps."""
    cfg = configs.scaled(configs.CUBE_256, [256, 16, 4])
    print_case("llg ddi", "256x16x4 row tiles")
    K = configs.site_anisotropy(256 * 16 * 4, 3)
    e = engine(cfg, site_K=K)
    S = oracle.System(cfg, site_K=K)
    s0 = oracle.random_spins(S.nos, 6)
    e.set_spins(s0)
    n = 30
    s, _, _ = _run_llg(e, 2, n)
    s_ref, _, _, _ = S.llg(s0, 2, n, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()


def test_llg_thermal_counter_rng(engine, oracle):
    cfg = configs.scaled(configs.FILM_128, [12, 10, 4])
    print_case("llg thermal philox")
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 8)
    e.set_spins(s0)
    n = 40
    s, _, _ = _run_llg(e, 3, n)
    # step 0 of a method is its first iteration; the seed is llg_seed
    s_ref, _, _, _ = S.llg(s0, 3, n, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=cfg["llg_temperature"],
                           seed=cfg["llg_seed"])
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()


def test_llg_thermal_injected_noise(engine, oracle):
    from spirit_amd import hip, simulation
    cfg = configs.scaled(configs.FILM_128, [10, 8, 3])
    print_case("llg thermal injected")
    e = engine(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 9)
    e.set_spins(s0)
    n = 25
    noise = np.random.default_rng(77).normal(size=(n, S.nos, 3))
    simulation.start(e.p, simulation.METHOD_LLG, 2, single_shot=True)
    hip.set_thermal_noise(e.p, noise)
    simulation.n_shot(e.p, n)
    s = e.spins()
    simulation.stop(e.p)
    s_ref, _, _, _ = S.llg(s0, 2, n, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=cfg["llg_temperature"],
                           noise=noise)
    assert np.abs(s - s_ref).max() <= 1e-9, np.abs(s - s_ref).max()


@pytest.mark.parametrize("solver", list(SOLVERS.keys()))
def test_larmor_precession_api(engine, solver):
    """core/test/test_physics.cpp:20-88 through the C API."""
    from spirit_amd import configuration, simulation, system
    from spirit_amd.parameters import llg
    L = GOLD["larmor"]
    print_case("larmor", solver)
    e = engine(L["cfg"])
    llg.set_damping(e.p, L["damping"])
    a = llg.get_damping(e.p)
    dt = llg.get_timestep(e.p)
    dtg = dt * 0.1760859644 / (1 + a * a)
    configuration.domain(e.p, [1, 0, 0])
    simulation.start(e.p, simulation.METHOD_LLG, SOLVERS[solver], single_shot=True)
    d = system.get_spin_directions(e.p)
    for i in range(L["steps"]):
        simulation.single_shot(e.p)
        phi = dtg * (i + 1) * 1.0
        sz = math.tanh(a * dtg * (i + 1) * 1.0)
        sx = math.cos(phi) * math.sqrt(1 - sz * sz)
        assert abs(d[0, 0] - sx) < L["tolerance"]
        assert abs(d[0, 2] - sz) < L["tolerance"]
    simulation.stop(e.p)


def test_anisotropy_closed_forms_api(engine):
    """core/test/test_anisotropy.cpp on the GPU path."""
    from spirit_amd import hamiltonian
    A = GOLD["anisotropy"]
    e = engine(A["cfg"])
    hamiltonian.set_anisotropy(e.p, A["K"], [0, 0, 1])
    K = float(np.float32(A["K"]))
    n = e.nos
    e.set_spins(np.tile([1.0, 0, 0], (n, 1)))
    _, e1 = e.gradient()
    e.set_spins(np.tile([0, 0, 1.0], (n, 1)))
    g, e2 = e.gradient()
    assert abs((e1 - e2) - K * n) < 1e-8
    np.testing.assert_allclose(g, np.tile([0, 0, -2 * K], (n, 1)), atol=1e-14)


def test_direct_minimisation_reaches_reference_minimum(engine):
    """core/test/test_solvers.cpp:40-72 for the LLG solvers of the HIP engine."""
    from spirit_amd import configuration, quantities, simulation, system
    from spirit_amd.parameters import llg
    G = GOLD["solvers"]
    e = engine(G["cfg"])
    llg.set_direct_minimization(e.p, True)
    for name in ["heun", "sib", "depondt", "rk4"]:
        print_case("minimisation", name)
        configuration.plus_z(e.p)
        sk = G["skyrmion"]
        configuration.skyrmion(e.p, sk["r"], sk["order"], sk["phase"], sk["upDown"], sk["achiral"], sk["rl"])
        simulation.start(e.p, simulation.METHOD_LLG, SOLVERS[name])
        E = system.get_energy(e.p)
        m = quantities.get_magnetization(e.p)
        assert abs(E - G["energy_minimum"]) < G["tolerance"], (name, E)
        assert abs(m[2] - G["mz_minimum"]) < G["tolerance"], (name, m)


def test_full_size_gradient_film(engine, oracle):
    """One force evaluation at the full 128x128x32 DDI film size against the oracle."""
    cfg = configs.FILM_128
    print_case("gradient", "film 128x128x32 full size")
    e = engine(cfg)
    S = oracle.System(cfg)
    s = oracle.random_spins(S.nos, 3)
    e.set_spins(s)
    g, E = e.gradient()
    g0, E0 = S.gradient_and_energy(s)
    assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    assert abs(E - E0) <= 1e-11 * abs(E0)


def test_headline_cube_against_oracle(engine, oracle):
    """The measured workload itself: 256x256x128 with the dipolar field (padded 512x512x256)
    and per-site K_i, the gradient and energy of a random configuration, then 2 Depondt steps
    (and 2 Heun, 2 SIB and 1 RK4 step from the same start), against the oracle (whose set-up
    and steps take about a minute on the host cores)."""
    cfg = configs.CUBE_256
    print_case("headline cube 256x256x128 against the oracle")
    K = configs.site_anisotropy(256 * 256 * 128, 20006)
    e = engine(cfg, site_K=K)
    S = oracle.System(cfg, site_K=K)
    s0 = oracle.random_spins(S.nos, 20006)
    e.set_spins(s0)
    g, E = e.gradient()
    g0, E0 = S.gradient_and_energy(s0)
    assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    assert abs(E - E0) <= 1e-11 * abs(E0)
    # the measured solver first, then the other three on the same workload (RK4: one step)
    for name, n in (("depondt", 2), ("heun", 2), ("sib", 2), ("rk4", 1)):
        e.set_spins(s0)
        s, _, _ = _run_llg(e, SOLVERS[name], n)
        s_ref, _, _, _ = S.llg(s0, SOLVERS[name], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
        err = float(np.abs(s - s_ref).max())
        print("  %s: max |s - s_oracle| after %d steps = %.3e" % (name, n, err), flush=True)
        assert err <= 1e-9, name


def test_full_size_properties_cube(engine):
    """256x256x128 with DDI (padded 512x512x256): the gradient is affine in the spins
    (g(a+b) = g(a) + g(b) - g(0)) and a few Depondt steps keep |s| = 1."""
    from spirit_amd import hip, simulation
    cfg = configs.CUBE_256
    print_case("properties", "cube 256x256x128")
    n = 256 * 256 * 128
    e = engine(cfg, site_K=configs.site_anisotropy(n, 5))
    a = configs.random_spins(n, 1)
    b = configs.random_spins(n, 2)
    e.set_spins(a)
    ga, _ = e.gradient()
    e.set_spins(b)
    gb, _ = e.gradient()
    e.set_spins(a + b)
    gab, _ = e.gradient()
    e.set_spins(np.zeros((n, 3)))
    g0, _ = e.gradient()
    err = np.abs(gab - (ga + gb - g0)).max() / np.abs(gab).max()
    assert err < 1e-12, err
    e.set_spins(a)
    simulation.start(e.p, simulation.METHOD_LLG, 2, single_shot=True)
    info = hip.ddi_info(e.p)
    print(info)
    assert "512x512x256" in info and "real, mirror-symmetric in b and c" in info
    simulation.n_shot(e.p, 3)
    s = e.spins()
    simulation.stop(e.p)
    assert np.abs(np.linalg.norm(s, axis=1) - 1).max() < 1e-13


@pytest.mark.parametrize("case,expect", [("film_bloch_ddi_open", "real, mirror-symmetric in b and c"),
                                         ("skewed_two_basis_ddi_periodic_images", "complex")])
def test_ddi_spectrum_format(engine, case, expect):
    """The kernel spectrum is stored compressed only where its symmetries hold: an open
    orthogonal film is even in r (real spectrum, mirror-symmetric); a skewed two-atom basis
    is not."""
    from spirit_amd import hip, simulation
    print_case("ddi format", case)
    e = engine(GRAD_CASES[case])
    simulation.start(e.p, simulation.METHOD_LLG, 2, single_shot=True)
    info = hip.ddi_info(e.p)
    simulation.stop(e.p)
    print(info)
    assert ("spectrum " + expect + ";") in info


ENERGY_TERMS = {"Zeeman": (0, 1.0), "Anisotropy": (1, 0.5), "Cubic anisotropy": (2, 0.25), "Exchange": (3, 0.5),
                "DMI": (4, 0.5), "DDI": (5, 0.5), "Quadruplets": (7, 0.25)}


@pytest.mark.parametrize("case", ["square_neel_periodic", "film_bloch_ddi_open", "hex_neel_periodic",
                                  "quadruplets_two_basis", "ddi_cutoff_periodic"])
def test_energy_contributions_per_spin(engine, oracle, case):
    """Energy_Contributions_per_Spin (Hamiltonian_Heisenberg.cpp:262-535) against the oracle's
    per-term gradients: a term homogeneous of degree n in s has E_i = (1/n) s_i . g_i (Zeeman
    n = 1, anisotropy / exchange / DMI / DDI n = 2, cubic n = 4).  The contributions add up to
    the total energy, and System_Get_Energy_Array returns their sums after System_Update_Data."""
    from spirit_amd import hip, system
    cfg = dict(GRAD_CASES[case])
    if case == "square_neel_periodic":
        cfg["cubic_anisotropy_magnitude"] = 0.3
    print_case("energy per spin", case)
    K = configs.site_anisotropy(int(np.prod(cfg["n_basis_cells"])), 4) if case == "film_bloch_ddi_open" else None
    e = engine(cfg, site_K=K)
    S = oracle.System(cfg, site_K=K)
    s = oracle.random_spins(S.nos, 12)
    e.set_spins(s)
    contrib = hip.energy_contributions_per_spin(e.p, S.nos)
    names = system.get_energy_array_names(e.p)
    assert list(contrib) == names and len(names) >= 3
    total = 0.0
    for name, E_i in contrib.items():
        term, scale = ENERGY_TERMS[name]
        g = S.gradient_term(term, s)
        ref = scale * np.einsum("ij,ij->i", s, g)
        assert np.abs(E_i - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), name
        total += E_i.sum()
    _, E0 = S.gradient_and_energy(s)
    assert abs(total - E0) <= 1e-11 * abs(E0)
    system.update_data(e.p)
    arr = system.get_energy_array(e.p, divide_by_nspins=False)
    np.testing.assert_allclose(arr, [float(np.float32(c.sum())) for c in contrib.values()], rtol=1e-6)
