"""Hamiltonian_Write_Hessian (host code, runs without a GPU): the embedding Hessian written
as triplets and as a dense matrix, checked against the Jacobian of the oracle's gradient
terms -- the check core/test/test_physics.cpp:95-141 makes of the dense Hessian with finite
differences. The gradients of exchange, DMI and uniaxial anisotropy are linear in the spins,
so the Jacobian is exact column by column: J[:, k] = grad(e_k).

As the reference's Sparse_Hessian (core/src/engine/Hamiltonian_Heisenberg.cpp:1288-1308),
the anisotropy block -2 K n n^T is written onto the diagonal as its row sums; exchange and
DMI entries equal the Jacobian. Values carry six significant digits (%g)."""
import json
import os

import numpy as np
import pytest

from spirit_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))

# simple cubic, mixed boundaries, two exchange shells, Bloch DMI, tilted uniaxial axis, field
SC = dict(bravais_lattice="sc", n_basis_cells=[4, 3, 2], boundary_conditions=[1, 0, 1], mu_s=[2.0],
          external_field_magnitude=0.5, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.2,
          anisotropy_normal=[0.6, 0.0, 0.8], n_shells_exchange=2, jij=[1.0, 0.25], n_shells_dmi=1, dij=[0.3],
          dm_chirality=1, ddi_method="none")
# skewed two-atom basis (core/test/input/physics_ddi.cfg's lattice) with exchange and Neel DMI
SKEW = dict(GOLD["physics_ddi"]["cfg"], hamiltonian="heisenberg_neighbours", n_basis_cells=[3, 2, 2],
            n_shells_exchange=2, jij=[1.0, 0.5], n_shells_dmi=1, dij=[0.7], dm_chirality=-1,
            anisotropy_magnitude=0.1, anisotropy_normal=[0, 0, 1], ddi_method="none")


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


def read_triplets(path):
    rows, cols, vals = [], [], []
    with open(path) as f:
        for line in f:
            r, c, v = line.rstrip("\n").split("\t")
            rows.append(int(r))
            cols.append(int(c))
            vals.append(float(v))
    return np.array(rows), np.array(cols), np.array(vals)


def jacobian(S, term):
    n = 3 * S.nos
    J = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        J[:, k] = S.gradient_term(term, e.reshape(-1, 3)).ravel()
    return J


def expected_hessian(S):
    """Exchange and DMI Jacobians plus the anisotropy block's row sums on the diagonal."""
    A = jacobian(S, 1)
    n = A.shape[0]
    rowsum = np.array([A[r, 3 * (r // 3):3 * (r // 3) + 3].sum() for r in range(n)])
    assert np.allclose(A - np.diag(np.diag(A)), np.where(np.arange(n)[:, None] // 3 == np.arange(n)[None, :] // 3,
                                                            A - np.diag(np.diag(A)), 0.0))
    return jacobian(S, 3) + jacobian(S, 4) + np.diag(rowsum)


@pytest.mark.parametrize("name,cfg", [("sc_mixed_bc", SC), ("skewed_basis", SKEW)])
def test_hessian_triplets_match_oracle_jacobian(make_state, tmp_path, oracle, name, cfg):
    from spirit_amd import hamiltonian
    p = make_state(cfg)
    S = oracle.System(cfg)
    n = 3 * S.nos
    path = str(tmp_path / "hessian.txt")
    hamiltonian.write_hessian(p, path, triplet_format=True)
    r, c, v = read_triplets(path)
    # column-major order, rows ascending within a column, no duplicates
    key = c * n + r
    assert np.all(np.diff(key) > 0)
    assert r.min() >= 0 and r.max() < n and c.min() >= 0 and c.max() < n
    H = np.zeros((n, n))
    H[r, c] = v
    want = expected_hessian(S)
    np.testing.assert_allclose(H, want, rtol=5e-6, atol=1e-12)
    # every non-zero of the expected matrix is an entry of the file
    nz = np.argwhere(np.abs(want) > 0)
    stored = set(zip(r.tolist(), c.tolist()))
    assert all((int(a), int(b)) in stored for a, b in nz)
    # exchange and DMI are symmetric; the anisotropy diagonal is too
    np.testing.assert_allclose(H, H.T, rtol=5e-6, atol=1e-12)


def test_hessian_dense_matches_triplets(make_state, tmp_path, oracle):
    from spirit_amd import hamiltonian
    p = make_state(SC)
    S = oracle.System(SC)
    n = 3 * S.nos
    tpath, dpath = str(tmp_path / "h_triplets.txt"), str(tmp_path / "h_dense.txt")
    hamiltonian.write_hessian(p, tpath, triplet_format=True)
    hamiltonian.write_hessian(p, dpath, triplet_format=False)
    r, c, v = read_triplets(tpath)
    text = open(dpath).read()
    lines = text.split("\n")
    assert lines[-1] == "" and len(lines) == n + 1
    D = np.zeros((n, n))
    for i, line in enumerate(lines[:-1]):
        # one "value " token per column, trailing space included
        assert line.endswith(" ")
        tok = line[:-1].split(" ")
        assert len(tok) == n
        D[i] = [float(t) for t in tok]
    H = np.zeros((n, n))
    H[r, c] = v
    np.testing.assert_array_equal(D, H)


def test_hessian_bad_file_is_reported_not_raised(make_state, tmp_path):
    from spirit_amd import hamiltonian
    p = make_state(SC)
    hamiltonian.write_hessian(p, str(tmp_path / "no_such_dir" / "h.txt"))
    assert not os.path.exists(tmp_path / "no_such_dir")


def _dipole_jacobian(pos, mu):
    """Jacobian of the point-dipole gradient: block (i, j) = -mu0 muB^2 mu_i mu_j / (4 pi |r|^3 1e-30)
    (3 r^ r^T - I) for i != j (positions in Angstrom), zero on the diagonal blocks."""
    pref = 2.0133545e-28 * 0.057883817555 ** 2 / (4 * np.pi * 1e-30)
    n = len(pos)
    J = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for j in range(n):
            if i != j:
                r = pos[j] - pos[i]
                d = np.linalg.norm(r)
                u = r / d
                J[3 * i:3 * i + 3, 3 * j:3 * j + 3] = -pref * mu[i] * mu[j] / d ** 3 * (3 * np.outer(u, u) - np.eye(3))
    return J


@pytest.mark.parametrize("name,cfg", [("sc_mixed_bc", SC), ("skewed_basis", SKEW)])
def test_oracle_dense_hessian(oracle, name, cfg):
    """The oracle's dense Hessian (Hamiltonian_Heisenberg::Hessian): exchange and DMI Jacobians
    plus the anisotropy row sums on the diagonal."""
    S = oracle.System(cfg)
    np.testing.assert_allclose(S.hessian(), expected_hessian(S), rtol=1e-14, atol=1e-14)


def test_oracle_dense_hessian_dipolar_block(oracle):
    """With a dipolar method set, the dense Hessian adds the point-dipole Jacobian over the
    positions without periodic images (Hamiltonian_Heisenberg.cpp:1210-1245), whatever the
    method and the boundary conditions."""
    for method, radius in (("fft", 0.0), ("cutoff", 2.0)):
        cfg = dict(SKEW, ddi_method=method, ddi_radius=radius, ddi_n_periodic_images=[1, 1, 1])
        S = oracle.System(cfg)
        base = oracle.System(dict(cfg, ddi_method="none"))
        mu = np.array([2.17] * S.nos)
        J = _dipole_jacobian(S.positions(), mu)
        np.testing.assert_allclose(S.hessian() - base.hessian(), J, rtol=0, atol=1e-12 * np.abs(J).max())
