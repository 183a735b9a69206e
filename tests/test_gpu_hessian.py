"""GPU embedding Hessian (Spirit_HIP_Hamiltonian_Get_Hessian) and its matrix-free product
(Spirit_HIP_Hamiltonian_Hessian_Vector_Product) against the oracle's dense Hessian
(Hamiltonian_Heisenberg::Hessian, Hamiltonian_Heisenberg.cpp:1119-1245), which the CPU tests
pin to the Jacobians of the gradient terms and to the point-dipole closed form.  Tolerance
1e-12 relative to the largest entry (the GPU contracts multiply-adds into FMAs)."""
import json
import os

import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))

SC = dict(bravais_lattice="sc", n_basis_cells=[4, 3, 2], boundary_conditions=[1, 0, 1], mu_s=[2.0],
          external_field_magnitude=0.5, anisotropy_magnitude=0.2, anisotropy_normal=[0.6, 0.0, 0.8],
          n_shells_exchange=2, jij=[1.0, 0.25], n_shells_dmi=1, dij=[0.3], dm_chirality=1, ddi_method="none")
SKEW = dict(GOLD["physics_ddi"]["cfg"], hamiltonian="heisenberg_neighbours", n_basis_cells=[3, 2, 2],
            n_shells_exchange=2, jij=[1.0, 0.5], n_shells_dmi=1, dij=[0.7], dm_chirality=-1,
            anisotropy_magnitude=0.1, anisotropy_normal=[0, 0, 1], ddi_method="none")
CASES = {
    "sc_mixed_bc": SC,
    "skewed_basis": SKEW,
    "skewed_basis_ddi_images": dict(SKEW, ddi_method="fft", ddi_n_periodic_images=[1, 1, 1]),
    "film_ddi_open": configs.scaled(configs.FILM_128, [12, 10, 4]),
    "hex_neel_cutoff": dict(bravais_lattice="hex2d", n_basis_cells=[9, 7, 1], boundary_conditions=[1, 1, 0],
                            mu_s=[2.0], anisotropy_magnitude=0.4, n_shells_exchange=1, jij=[10.0], n_shells_dmi=1,
                            dij=[6.0], dm_chirality=2, ddi_method="cutoff", ddi_radius=2.0),
}


def _state(tmp_path, cfg, name="input.cfg"):
    from spirit_amd import state
    path = str(tmp_path / name)
    configs.write_config(cfg, path)
    p = state.setup(path, True)
    assert p
    return p


@pytest.mark.parametrize("name", list(CASES))
def test_dense_hessian_matches_oracle(tmp_path, oracle, name):
    from spirit_amd import hip, state, system
    print("CASE hessian", name, flush=True)
    cfg = CASES[name]
    p = _state(tmp_path, cfg)
    try:
        nos = system.get_nos(p)
        S = oracle.System(cfg)
        H0 = S.hessian()
        H = hip.hessian(p, nos)
        scale = np.abs(H0).max()
        assert np.abs(H - H0).max() <= 1e-12 * scale, np.abs(H - H0).max() / scale
        v = oracle.random_spins(nos, 3) * np.linspace(0.5, 1.5, nos)[:, None]
        hv = hip.hessian_vector_product(p, v)
        want = (H0 @ v.ravel()).reshape(-1, 3)
        assert np.abs(hv - want).max() <= 1e-12 * np.abs(want).max()
    finally:
        state.delete(p)


def test_dense_hessian_site_anisotropy(tmp_path, oracle):
    """Per-site K_i along one axis enters the diagonal as the uniaxial term's row sums."""
    from spirit_amd import hip, state, system
    print("CASE hessian per-site K", flush=True)
    cfg = configs.scaled(configs.CUBE_256, [6, 5, 3])
    p = _state(tmp_path, cfg)
    try:
        nos = system.get_nos(p)
        K = configs.site_anisotropy(nos, 9)
        hip.set_site_anisotropy(p, K, (0.0, 0.0, 1.0))
        S = oracle.System(cfg, site_K=K)
        H0 = S.hessian()
        H = hip.hessian(p, nos)
        assert np.abs(H - H0).max() <= 1e-12 * np.abs(H0).max()
    finally:
        state.delete(p)


def test_hessian_row_chunks_and_product(tmp_path):
    """A 64x56 film with dipoles (3584 spins, a 10752^2 matrix written in several row launches):
    symmetric, and the matrix-free product equals the dense matrix times v."""
    from spirit_amd import hip, state, system
    print("CASE hessian 64x56", flush=True)
    cfg = configs.scaled(configs.FILM_128, [64, 56, 1])
    p = _state(tmp_path, cfg)
    try:
        nos = system.get_nos(p)
        H = hip.hessian(p, nos)
        assert np.abs(H - H.T).max() <= 1e-13 * np.abs(H).max()
        v = configs.random_spins(nos, 4)
        hv = hip.hessian_vector_product(p, v)
        want = (H @ v.ravel()).reshape(-1, 3)
        assert np.abs(hv - want).max() <= 1e-12 * np.abs(want).max()
    finally:
        state.delete(p)
