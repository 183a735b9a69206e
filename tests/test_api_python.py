"""Python API surface beyond the hot path, written after the reference's own Python tests
(core/python/test/quantities.py, system.py): the topological charge and its density per
triangle, and the energy contributions as a dictionary."""
import math

import numpy as np
import pytest

from spirit_amd import configs


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


@pytest.mark.parametrize("lattice", ["sc", "hex2d"])
def test_topological_charge_density(make_state, lattice):
    """A skyrmion (radius 5, core down in an up background, as the reference's quantities test
    sets it) on a periodic 40x40 lattice carries Q = -1; the per-triangle densities sum to Q,
    there are two triangles per cell and each names three spins of the lattice."""
    from spirit_amd import configuration, quantities
    p = make_state(dict(bravais_lattice=lattice, n_basis_cells=[40, 40, 1], boundary_conditions=[1, 1, 0],
                        mu_s=[1.0], n_shells_exchange=0, n_shells_dmi=0))
    configuration.plus_z(p)
    assert abs(quantities.get_topological_charge(p)) < 1e-6
    configuration.skyrmion(p, 5.0, pos=[1.5, 0.0, 0.0])
    Q = quantities.get_topological_charge(p)
    dens, tri = quantities.get_topological_charge_density(p)
    assert abs(Q + 1.0) < 1e-4, Q
    assert abs(sum(dens) - Q) < 1e-4
    assert len(dens) == len(tri) == 2 * 40 * 40
    assert all(0 <= i < 1600 for t in tri for i in t)
    assert max(abs(d) for d in dens) < 0.5  # no triangle spans a half sphere


def test_topological_charge_outside_2d_one_atom(make_state):
    """Three-dimensional lattices (and several basis atoms) have no density here: Q = 0 and
    empty lists."""
    from spirit_amd import configuration, quantities
    p = make_state(dict(bravais_lattice="sc", n_basis_cells=[8, 8, 3], boundary_conditions=[1, 1, 0], mu_s=[1.0],
                        n_shells_exchange=0, n_shells_dmi=0))
    configuration.random(p)
    assert quantities.get_topological_charge(p) == 0.0
    assert quantities.get_topological_charge_density(p) == ([], [])


@pytest.mark.gpu
def test_energy_contributions_dictionary(make_state):
    """system.get_energy_contributions after update_data: one entry per contribution of the
    Hamiltonian, summing to the total energy (the reference's system test, on a 2x2 lattice
    with exchange, DMI and a domain)."""
    from spirit_amd import configuration, system
    print("CASE energy contributions dictionary", flush=True)
    cfg = dict(bravais_lattice="sc", n_basis_cells=[2, 2, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=0.0, anisotropy_magnitude=0.0, n_shells_exchange=1, jij=[10.0],
               n_shells_dmi=1, dij=[6.0], dm_chirality=1)
    p = make_state(cfg)
    configuration.plus_z(p)
    configuration.domain(p, [0, 0, -1], border_cylindrical=2)
    system.update_data(p)
    E_contribs = system.get_energy_contributions(p, divide_by_nspins=False)
    E = system.get_energy(p)
    assert set(E_contribs) == {"Exchange", "DMI"}, E_contribs
    assert abs(sum(E_contribs.values()) - E) <= 1e-5 * max(1.0, abs(E))


def test_dw_skyrmion_and_hopfion(make_state):
    """The configurations the reference's configuration test sets (core/python/test/configuration.py):
    a domain-wall skyrmion (core down, background up, the wall where s_z changes sign at about
    dw_radius) and a hopfion (unit spins, up along the torus axis, a closed tube of down spins
    through the torus ring)."""
    from spirit_amd import configuration, geometry, system
    p = make_state(dict(bravais_lattice="sc", n_basis_cells=[30, 30, 1], boundary_conditions=[1, 1, 0], mu_s=[1.0],
                        n_shells_exchange=0, n_shells_dmi=0))
    configuration.plus_z(p)
    configuration.dw_skyrmion(p, 5.0, 2.0)
    s = np.array(system.get_spin_directions(p))
    pos = np.array(geometry.get_positions(p))
    c = np.array(geometry.get_center(p))
    r = np.hypot(pos[:, 0] - c[0], pos[:, 1] - c[1])
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-12)
    assert s[np.argmin(r), 2] < -0.99
    assert np.all(s[(r > 8) & (r < 14), 2] > 0.9)
    inside, outside = s[(r > 3.5) & (r < 4.5), 2], s[(r > 5.5) & (r < 6.5), 2]
    assert inside.max() < 0 < outside.min()
    q = make_state(dict(bravais_lattice="sc", n_basis_cells=[24, 24, 24], boundary_conditions=[0, 0, 0],
                        mu_s=[1.0], n_shells_exchange=0, n_shells_dmi=0))
    configuration.plus_z(q)
    configuration.hopfion(q, 5.0)
    s = np.array(system.get_spin_directions(q))
    pos = np.array(geometry.get_positions(q))
    c = np.array(geometry.get_center(q))
    d = np.linalg.norm(pos - c, axis=1)
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-12)
    assert s[:, 2].min() < -0.9  # the down tube
    rho = np.hypot(pos[:, 0] - c[0], pos[:, 1] - c[1])
    ring = (np.abs(pos[:, 2] - c[2]) < 1.0) & (np.abs(rho - 5.0) < 1.0)
    assert s[ring, 2].mean() < -0.5
    axis = (rho < 1.0) & (d < 5.0 * math.pi)
    assert axis.sum() > 10 and s[axis, 2].min() > 0.9  # up along the torus axis (sites 0.7 off it)
