"""GPU parity of the spin-transfer torque (monolayer and gradient, Method_LLG.cpp:150-212) and
of the temperature gradient (Method_LLG.cpp:66-110, Vectormath.cpp:633-652): the engine
through the C API against the oracle on the same seeded inputs.  Spins after N steps to
1e-9 max-abs per component (FP64 on both sides; the GPU contracts multiply-adds into FMAs
and weights the gradient's three lattice differences by B^-1 j instead of forming J)."""
import math

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

SOLVERS = {"depondt": 2, "heun": 3, "sib": 1, "rk4": 4}
GAMMA, MU_B = 0.1760859644, 0.057883817555


def print_case(*a):
    print("CASE", *a, flush=True)


def film(n, bc, **kw):
    cfg = dict(bravais_lattice="sc", n_basis_cells=n, boundary_conditions=bc, mu_s=[2.0],
               external_field_magnitude=0.5, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.05,
               anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[5.0], n_shells_dmi=1, dij=[1.5],
               dm_chirality=2, ddi_method="none", llg_dt=1e-3, llg_damping=0.25, llg_temperature=0.0,
               llg_seed=31, llg_force_convergence=1e-12)
    cfg.update(kw)
    return cfg


@pytest.fixture
def make_state(tmp_path):
    from spirit_amd import state as st
    made = []

    def make(cfg, name="input.cfg"):
        path = str(tmp_path / name)
        configs.write_config(cfg, path)
        p = st.setup(path, True)
        assert p
        made.append(p)
        return p

    yield make
    for p in made:
        st.delete(p)


def spins_of(p):
    from spirit_amd import system
    return np.array(system.get_spin_directions(p))


def set_spins(p, s):
    from spirit_amd import system
    system.get_spin_directions(p)[:] = s


def run_n_shot(p, solver, n):
    from spirit_amd import simulation, system
    simulation.start(p, simulation.METHOD_LLG, solver, single_shot=True)
    simulation.n_shot(p, n)
    s = spins_of(p)
    E = system.get_energy(p)
    T = simulation.get_max_torque_norm(p)
    simulation.stop(p)
    return s, E, T


STT_MONO = dict(llg_beta=0.1, llg_stt_magnitude=0.3, llg_stt_use_gradient=0,
                llg_stt_polarisation_normal=[0.3, 0.4, 0.866])
STT_GRAD = dict(llg_beta=0.05, llg_stt_magnitude=0.6, llg_stt_use_gradient=1,
                llg_stt_polarisation_normal=[1.0, 0.5, 0.2])


@pytest.mark.parametrize("mode", ["monolayer", "gradient"])
@pytest.mark.parametrize("solver", list(SOLVERS))
def test_stt_trajectory(make_state, oracle, mode, solver):
    """A 20x16x3 film, periodic along a only (the gradient takes one-sided differences at the
    open b and c faces), exchange, Neel DMI, anisotropy and field: 200 steps with the torque
    set in the input file; spins, energy and max torque against the oracle."""
    extra = STT_MONO if mode == "monolayer" else STT_GRAD
    cfg = film([20, 16, 3], [1, 0, 0], **extra)
    print_case("stt", mode, solver)
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 41)
    set_spins(p, s0)
    n = 200
    s, E, T = run_n_shot(p, SOLVERS[solver], n)
    stt = (extra["llg_stt_magnitude"], extra["llg_stt_polarisation_normal"], bool(extra["llg_stt_use_gradient"]))
    s_ref, E_ref, T_ref, _ = S.llg(s0, SOLVERS[solver], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"],
                                   beta=extra["llg_beta"], stt=stt)
    err = float(np.abs(s - s_ref).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9
    assert abs(E - E_ref) <= 1e-5 * abs(E_ref) + 1e-3  # System_Get_Energy returns a float
    assert abs(T - T_ref) <= 1e-5 * T_ref + 1e-7
    # the torque matters: without it the run ends elsewhere
    s_plain, _, _, _ = S.llg(s0, SOLVERS[solver], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"])
    assert np.abs(s_plain - s_ref).max() > 1e-3


def test_stt_gradient_skewed_two_basis(make_state, oracle):
    """The gradient torque on a skewed two-atom lattice (explicit Bravais vectors, lattice
    constant 1.7, open along c): the differences along the three translations weighted by
    B^-1 j must equal the oracle's J = M B^-1 applied to j.  Heun, 150 steps."""
    cfg = dict(bravais_vectors=[[1.0, 0.0, 0.0], [0.5, 0.8, 0.0], [0.2, 0.3, 0.9]], n_basis_cells=[12, 10, 3],
               basis=[[0.0, 0.0, 0.0], [0.3, 0.4, 0.2]], lattice_constant=1.7, boundary_conditions=[1, 1, 0],
               mu_s=[1.0, 2.0], external_field_magnitude=1.0, external_field_normal=[0, 0, 1], n_shells_exchange=1,
               jij=[4.0], n_shells_dmi=0, ddi_method="none", llg_dt=1e-3, llg_damping=0.3, llg_temperature=0.0,
               llg_seed=3, llg_force_convergence=1e-12, llg_beta=0.2, llg_stt_magnitude=0.8, llg_stt_use_gradient=1,
               llg_stt_polarisation_normal=[0.3, -0.5, 0.81])
    print_case("stt gradient skewed two-atom basis")
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 42)
    set_spins(p, s0)
    s, _, _ = run_n_shot(p, SOLVERS["heun"], 150)
    s_ref, _, _, _ = S.llg(s0, SOLVERS["heun"], 150, dt=cfg["llg_dt"], damping=cfg["llg_damping"], beta=0.2,
                           stt=(0.8, [0.3, -0.5, 0.81], True))
    err = float(np.abs(s - s_ref).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9


def test_stt_gradient_helix_translation(make_state):
    """With beta = alpha the gradient torque turns every spin of a helix about its axis by
    -dt gamma a_j / mu_B * sin(k a) / a per step (the helix moves rigidly along the current):
    the engine's spins after 400 Depondt steps against that closed form."""
    N, m = 48, 5
    cfg = dict(bravais_lattice="sc", n_basis_cells=[N, 1, 1], boundary_conditions=[1, 1, 1], mu_s=[2.0],
               n_shells_exchange=0, n_shells_dmi=0, ddi_method="none", llg_dt=1e-3, llg_damping=0.2,
               llg_temperature=0.0, llg_seed=1, llg_force_convergence=1e-12, llg_beta=0.2, llg_stt_magnitude=0.5,
               llg_stt_use_gradient=1, llg_stt_polarisation_normal=[1.0, 0.0, 0.0])
    print_case("stt helix")
    p = make_state(cfg)
    k = 2 * math.pi * m / N
    x = np.arange(N, dtype=float)
    set_spins(p, np.stack([np.cos(k * x), np.sin(k * x), 0 * x], axis=1))
    n = 400
    s, _, _ = run_n_shot(p, SOLVERS["depondt"], n)
    ph = k * x - n * 1e-3 * GAMMA / MU_B * 0.5 * math.sin(k)
    want = np.stack([np.cos(ph), np.sin(ph), 0 * x], axis=1)
    err = float(np.abs(s - want).max())
    print("  max |s - closed form| = %.3e" % err, flush=True)
    assert err <= 1e-10


def test_stt_set_through_api_while_running(make_state, oracle):
    """Parameters_LLG_Set_STT and _Set_Non_Adiabatic_Damping on an image whose single-shot
    method is running: 60 steps without, then 120 with the gradient torque (the running
    method takes the new parameters), then 60 with the monolayer torque."""
    from spirit_amd import simulation
    from spirit_amd.parameters import llg
    cfg = film([16, 12, 2], [1, 1, 0])
    print_case("stt via api")
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 43)
    set_spins(p, s0)
    simulation.start(p, simulation.METHOD_LLG, SOLVERS["depondt"], single_shot=True)
    simulation.n_shot(p, 60)
    llg.set_non_adiabatic_damping(p, 0.15)
    llg.set_stt(p, True, 0.7, [0.0, 1.0, 0.0])
    assert llg.get_stt(p) == (pytest.approx(0.7), [0.0, 1.0, 0.0], True)
    assert llg.get_non_adiabatic_damping(p) == pytest.approx(0.15)
    simulation.n_shot(p, 120)
    llg.set_stt(p, False, 0.4, [0.0, 0.0, 0.0])  # a zero normal becomes (0, 0, 1), as the reference's
    assert llg.get_stt(p)[1] == [0.0, 0.0, 1.0]
    simulation.n_shot(p, 60)
    s = spins_of(p)
    simulation.stop(p)
    dt, al = cfg["llg_dt"], cfg["llg_damping"]
    r, _, _, _ = S.llg(s0, 2, 60, dt=dt, damping=al)
    r, _, _, _ = S.llg(r, 2, 120, dt=dt, damping=al, beta=float(np.float32(0.15)),
                       stt=(float(np.float32(0.7)), [0, 1, 0], True), step0=60)
    r, _, _, _ = S.llg(r, 2, 60, dt=dt, damping=al, beta=float(np.float32(0.15)),
                       stt=(float(np.float32(0.4)), [0, 0, 1], False), step0=180)
    err = float(np.abs(s - r).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("solver", ["heun", "depondt"])
def test_temperature_gradient_counter_rng(make_state, oracle, solver):
    """A temperature gradient along (1, 1, 0) from 5 K (3 K per length unit) on a 14x10x3 film
    with the dipolar field: each site's thermal amplitude from its own T_i, the counter-based
    thermal stream of the seed; 60 steps against the oracle."""
    cfg = configs.scaled(configs.FILM_128, [14, 10, 3], llg_temperature=5.0,
                         llg_temperature_gradient_inclination=3.0, llg_temperature_gradient_direction=[1.0, 1.0, 0.0])
    print_case("temperature gradient philox", solver)
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 44)
    set_spins(p, s0)
    n = 60
    s, _, _ = run_n_shot(p, SOLVERS[solver], n)
    s_ref, _, _, _ = S.llg(s0, SOLVERS[solver], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=5.0,
                           seed=cfg["llg_seed"], temperature_gradient=(3.0, [1.0, 1.0, 0.0]))
    err = float(np.abs(s - s_ref).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9
    s_hom, _, _, _ = S.llg(s0, SOLVERS[solver], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=5.0,
                           seed=cfg["llg_seed"])
    assert np.abs(s_hom - s_ref).max() > 1e-3


@pytest.mark.parametrize("T0,incl,direction", [(0.0, 2.0, [0.0, -1.0, 0.0]), (5.0, -2.0, [0.0, 1.0, 0.0])])
def test_temperature_gradient_injected_noise(make_state, oracle, T0, incl, direction):
    """At T = 0 a gradient alone switches the thermal field on (Method_LLG.cpp:72): 22 K at
    b = 0 falling to 0 K at b = 11; a negative inclination is cut at 0 K (Vectormath.cpp:651):
    5 K at b = 0, 0 K from b = 3 on.  Injected noise, SIB, 30 steps, set through
    Parameters_LLG_Set_Temperature and _Set_Temperature_Gradient."""
    from spirit_amd import hip, simulation
    from spirit_amd.parameters import llg
    cfg = film([16, 12, 1], [1, 1, 0])
    print_case("temperature gradient injected", T0, incl, direction)
    p = make_state(cfg)
    llg.set_temperature(p, T0, incl, direction)
    assert llg.get_temperature(p) == (T0, incl, direction)
    S = oracle.System(cfg)
    Ti = S.temperature_distribution(direction, T0, incl)
    assert Ti.min() == 0.0 and Ti.max() > 4.0
    s0 = oracle.random_spins(S.nos, 45)
    set_spins(p, s0)
    n = 30
    noise = np.random.default_rng(78).normal(size=(n, S.nos, 3))
    simulation.start(p, simulation.METHOD_LLG, SOLVERS["sib"], single_shot=True)
    hip.set_thermal_noise(p, noise)
    simulation.n_shot(p, n)
    s = spins_of(p)
    simulation.stop(p)
    s_ref, _, _, _ = S.llg(s0, SOLVERS["sib"], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=T0,
                           noise=noise, temperature_gradient=(incl, direction))
    err = float(np.abs(s - s_ref).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9


def test_batch_with_stt_and_gradient_replicas(make_state, oracle):
    """A batched run (Spirit_HIP_Simulation_LLG_Start_Batch) of three replicas, one with the
    gradient torque, one with the monolayer torque and a temperature gradient, one plain: each
    equals its oracle run (Depondt, 150 steps)."""
    from spirit_amd import hip
    cfgs = [film([18, 14, 2], [1, 0, 0], **STT_GRAD),
            film([10, 10, 1], [1, 1, 0], llg_temperature=2.0, llg_temperature_gradient_inclination=1.0,
                 llg_temperature_gradient_direction=[0.0, 1.0, 0.0], **STT_MONO),
            film([24, 8, 1], [0, 1, 0])]
    print_case("batch stt")
    ps = [make_state(c, "b%d.cfg" % k) for k, c in enumerate(cfgs)]
    s0s = []
    for k, p in enumerate(ps):
        from spirit_amd import system
        s0 = oracle.random_spins(system.get_nos(p), 60 + k)
        set_spins(p, s0)
        s0s.append(s0)
    infos = hip.llg_start_batch(ps, SOLVERS["depondt"], n_iterations=150)
    assert [i.total_iterations for i in infos] == [150] * 3
    for k, cfg in enumerate(cfgs):
        S = oracle.System(cfg)
        kw = {}
        if cfg.get("llg_stt_magnitude", 0) > 0:
            kw["beta"] = cfg["llg_beta"]
            kw["stt"] = (cfg["llg_stt_magnitude"], cfg["llg_stt_polarisation_normal"],
                         bool(cfg["llg_stt_use_gradient"]))
        if cfg.get("llg_temperature_gradient_inclination", 0) != 0:
            kw["temperature_gradient"] = (cfg["llg_temperature_gradient_inclination"],
                                          cfg["llg_temperature_gradient_direction"])
        r, _, _, _ = S.llg(s0s[k], SOLVERS["depondt"], 150, dt=cfg["llg_dt"], damping=cfg["llg_damping"],
                           temperature=cfg["llg_temperature"], seed=cfg["llg_seed"], **kw)
        err = float(np.abs(spins_of(ps[k]) - r).max())
        print("  replica %d: max |s - s_oracle| = %.3e" % (k, err), flush=True)
        assert err <= 1e-9


def test_stt_gradient_full_size_cube(make_state, oracle):
    """The gradient torque at the full 256x256x128 size (without the dipolar field, whose
    oracle run would take minutes): the stage kernels' STT instantiation over 32-row tiles of
    8.4M spins, 3 Depondt steps against the oracle."""
    cfg = dict(configs.CUBE_256, ddi_method="none", llg_beta=0.1, llg_stt_magnitude=0.5, llg_stt_use_gradient=1,
               llg_stt_polarisation_normal=[0.6, 0.8, 0.0])
    print_case("stt gradient 256x256x128")
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 46)
    set_spins(p, s0)
    s, _, _ = run_n_shot(p, SOLVERS["depondt"], 3)
    s_ref, _, _, _ = S.llg(s0, SOLVERS["depondt"], 3, dt=cfg["llg_dt"], damping=cfg["llg_damping"], beta=0.1,
                           stt=(0.5, [0.6, 0.8, 0.0], True))
    err = float(np.abs(s - s_ref).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9


def test_temperature_gradient_rk4_with_stt(make_state, oracle):
    """RK4 (the thermal field drawn once per step and read by the later stages) with a
    temperature gradient and the monolayer torque together, injected noise, 20 steps."""
    from spirit_amd import hip, simulation
    cfg = film([12, 10, 2], [1, 1, 0], llg_temperature=4.0, llg_temperature_gradient_inclination=1.5,
               llg_temperature_gradient_direction=[1.0, 0.0, 0.0], **STT_MONO)
    print_case("temperature gradient + stt, rk4")
    p = make_state(cfg)
    S = oracle.System(cfg)
    s0 = oracle.random_spins(S.nos, 47)
    set_spins(p, s0)
    n = 20
    noise = np.random.default_rng(79).normal(size=(n, S.nos, 3))
    simulation.start(p, simulation.METHOD_LLG, SOLVERS["rk4"], single_shot=True)
    hip.set_thermal_noise(p, noise)
    simulation.n_shot(p, n)
    s = spins_of(p)
    simulation.stop(p)
    s_ref, _, _, _ = S.llg(s0, SOLVERS["rk4"], n, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=4.0,
                           noise=noise, beta=STT_MONO["llg_beta"],
                           stt=(STT_MONO["llg_stt_magnitude"], STT_MONO["llg_stt_polarisation_normal"], False),
                           temperature_gradient=(1.5, [1.0, 0.0, 0.0]))
    err = float(np.abs(s - s_ref).max())
    print("  max |s - s_oracle| = %.3e" % err, flush=True)
    assert err <= 1e-9
