"""Batched ensembles of independent LLG replicas (Spirit_HIP_Simulation_LLG_Start_Batch):
uneven lattice sizes, per-replica fields, temperatures and seeds advance in one launch per
solver stage and must give exactly what running Simulation_LLG_Start on each replica in turn
gives (same kernels, same per-replica inputs, same thermal-noise counters: bit-identical)."""
import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu


def print_case(*a):
    print("CASE", *a, flush=True)


def replica_cfg(L, B, T, seed, **kw):
    cfg = dict(bravais_lattice="sc", n_basis_cells=[L, L, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=B, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.1,
               anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[10.0], n_shells_dmi=1, dij=[3.0],
               dm_chirality=2, ddi_method="none", llg_dt=1e-3, llg_damping=0.2, llg_temperature=T,
               llg_seed=seed, llg_force_convergence=1e-12)
    cfg.update(kw)
    return cfg


@pytest.fixture
def states(tmp_path):
    from spirit_amd import state as st
    made = []

    def make(cfgs, spins_seed=7):
        from spirit_amd import system
        out = []
        for k, cfg in enumerate(cfgs):
            path = str(tmp_path / ("r%d.cfg" % k))
            configs.write_config(cfg, path)
            p = st.setup(path, True)
            assert p
            n = system.get_nos(p)
            system.get_spin_directions(p)[:] = configs.random_spins(n, spins_seed + k)
            out.append(p)
            made.append(p)
        return out

    yield make
    for p in made:
        st.delete(p)


def spins(p):
    from spirit_amd import system
    return np.array(system.get_spin_directions(p))


@pytest.mark.parametrize("solver", [2, 3, 1, 4])
def test_batch_equals_sequential(states, solver):
    """Uneven replicas (L = 8..40), B and T per replica (thermal noise from each replica's seed)."""
    from spirit_amd import hip, simulation, system
    print_case("batch vs sequential", solver)
    rng = np.random.default_rng(11)
    Ls = [8, 40, 12, 24, 16, 32]
    cfgs = [replica_cfg(L, float(rng.uniform(0, 2)), float(rng.uniform(1, 50)), 100 + k) for k, L in enumerate(Ls)]
    seq = states(cfgs)
    for p in seq:
        simulation.start(p, simulation.METHOD_LLG, solver, n_iterations=150)
    bat = states(cfgs)
    infos = hip.llg_start_batch(bat, solver, n_iterations=150)
    for k in range(len(Ls)):
        np.testing.assert_array_equal(spins(bat[k]), spins(seq[k]))
        assert system.get_energy(bat[k]) == system.get_energy(seq[k])
        assert infos[k].total_iterations == 150


def test_batch_per_replica_convergence(states):
    """Direct minimisation at T=0: each replica stops at its own convergence iteration, the same
    one a lone Simulation_LLG_Start reaches (convergence tested every step)."""
    from spirit_amd import hip, simulation
    from spirit_amd.parameters import llg
    print_case("batch", "per-replica convergence")
    cfgs = [replica_cfg(L, B, 0.0, 1, llg_force_convergence=c)
            for L, B, c in [(8, 1.0, 1e-6), (12, 0.5, 1e-4), (10, 1.5, 1e-8)]]
    seq = states(cfgs)
    n_seq = []
    for p in seq:
        llg.set_direct_minimization(p, True)
        info = simulation.start(p, simulation.METHOD_LLG, 2, n_iterations=20000)
        n_seq.append(info.total_iterations)
    bat = states(cfgs)
    for p in bat:
        llg.set_direct_minimization(p, True)
    infos = hip.llg_start_batch(bat, 2, n_iterations=20000)
    assert [i.total_iterations for i in infos] == n_seq
    assert len(set(n_seq)) == 3 and max(n_seq) < 20000
    for k in range(3):
        np.testing.assert_array_equal(spins(bat[k]), spins(seq[k]))


def test_batch_with_dipolar_replicas(states):
    """Replicas with their own dipolar FFT plans (open films of different sizes, Heun at T > 0)
    next to a replica without DDI: the batched run equals the sequential runs."""
    from spirit_amd import hip, simulation
    print_case("batch", "ddi replicas")
    film = configs.scaled(configs.FILM_128, [12, 10, 4])
    cfgs = [film, configs.scaled(configs.FILM_128, [16, 8, 3], llg_seed=77),
            replica_cfg(20, 1.0, 30.0, 5)]
    seq = states(cfgs)
    for p in seq:
        simulation.start(p, simulation.METHOD_LLG, 3, n_iterations=60)
    bat = states(cfgs)
    hip.llg_start_batch(bat, 3, n_iterations=60)
    for k in range(len(cfgs)):
        np.testing.assert_array_equal(spins(bat[k]), spins(seq[k]))


def test_ensemble_replicas_match_oracle(states):
    """The cfg-4 uneven ensemble (seeded sides, B, T and seeds) against the oracle: one replica
    of each side 32, 48, 64 and 256 taken from configs.ensemble(256), advanced together by one
    batched Depondt run of 200 steps at their temperatures; each equals the oracle's run of
    that replica on the same counter-based thermal stream (its seed, steps 0..199) to 1e-9."""
    import oracle as O
    from spirit_amd import hip
    print_case("ensemble256 replicas vs oracle")
    ens = configs.ensemble(256)
    pick = []
    for L in (32, 48, 64, 256):
        pick.append(next(c for c in ens if c["n_basis_cells"][0] == L))
    assert all(c["llg_temperature"] > 0 for c in pick)
    bat = states(pick, spins_seed=500)
    infos = hip.llg_start_batch(bat, 2, n_iterations=200)
    assert [i.total_iterations for i in infos] == [200] * 4
    for k, cfg in enumerate(pick):
        n = cfg["n_basis_cells"][0] ** 2
        s0 = configs.random_spins(n, 500 + k)
        ref, _, _, _ = O.System(cfg).llg(s0, O.SOLVER_DEPONDT, 200, dt=cfg["llg_dt"], damping=cfg["llg_damping"],
                                        temperature=cfg["llg_temperature"], seed=cfg["llg_seed"])
        err = float(np.abs(spins(bat[k]) - ref).max())
        print("  L=%d B=%.3f T=%.2f seed=%d: max |s - s_oracle| = %.3e" % (
            cfg["n_basis_cells"][0], cfg["external_field_magnitude"], cfg["llg_temperature"], cfg["llg_seed"], err),
            flush=True)
        assert err < 1e-9


def test_ensemble256_full_properties(states):
    """The full 256-replica ensemble (about 4.8 million spins) through one batched Depondt run
    of 500 steps: every replica completes, every spin stays a unit vector to 1e-12, and the
    replicas of a sample (the first of each side) are bit-identical to their lone runs."""
    from spirit_amd import hip, simulation
    print_case("ensemble256 full batch")
    ens = configs.ensemble(256)
    bat = states(ens, spins_seed=900)
    infos = hip.llg_start_batch(bat, 2, n_iterations=500)
    assert all(i.total_iterations == 500 for i in infos)
    worst = 0.0
    for p in bat:
        s = spins(p)
        worst = max(worst, float(np.abs(np.linalg.norm(s, axis=1) - 1).max()))
    assert worst < 1e-12, worst
    seen = set()
    for k, cfg in enumerate(ens):
        L = cfg["n_basis_cells"][0]
        if L in seen:
            continue
        seen.add(L)
        lone = states([cfg], spins_seed=900 + k)[0]
        simulation.start(lone, simulation.METHOD_LLG, 2, n_iterations=500)
        np.testing.assert_array_equal(spins(lone), spins(bat[k]))
    assert len(seen) == 7
