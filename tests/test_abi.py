"""The C-ABI library: loads, exports every declared entry point, and its host-side
set-up (input parsing, neighbour tables) agrees with the oracle.  No GPU is used."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "spirit_amd", "libSpirit.so")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include", "Spirit")
    for f in os.listdir(inc):
        txt = open(os.path.join(inc, f)).read()
        for m in re.finditer(r"SPIRIT_API\s+[\w\s\*]*?\b(\w+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libSpirit.so not built")
    return ctypes.CDLL(LIB)


def test_headers_declare_the_reference_entry_points():
    names = declared_symbols()
    for must in ["State_Setup", "State_Delete", "Simulation_LLG_Start", "Simulation_GNEB_Start", "Simulation_N_Shot",
                 "Hamiltonian_Set_Field", "Hamiltonian_Set_DDI", "Configuration_Random", "System_Get_Spin_Directions",
                 "Parameters_LLG_Set_Time_Step", "Chain_Insert_Image_After", "Transition_Homogeneous",
                 "Quantity_Get_Magnetization"]:
        assert must in names
    assert len(names) > 150


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, missing


def _state(tmp_path, cfg):
    from spirit_amd import configs, state
    p = str(tmp_path / "in.cfg")
    configs.write_config(cfg, p)
    return state.setup(p, True)


@pytest.mark.parametrize("name", list(GOLD["neighbour_tables"].keys()))
def test_host_neighbour_tables_bit_exact(lib, tmp_path, name):
    from spirit_amd import hamiltonian, state
    case = GOLD["neighbour_tables"][name]
    cfg = dict(case["cfg"], n_shells_exchange=1, jij=[1.0], mu_s=[2.0])
    p = _state(tmp_path, cfg)
    try:
        idx, tr, J = hamiltonian.get_exchange_pairs(p)
        got = [[int(a), int(b)] + [int(x) for x in t] for (a, b), t in zip(idx, tr)]
        assert got == case["pairs"]
    finally:
        state.delete(p)


def test_host_pair_table_matches_oracle_skewed_basis(lib, tmp_path, oracle):
    from spirit_amd import hamiltonian, state
    cfg = dict(GOLD["physics_ddi"]["cfg"])
    cfg.update(hamiltonian="heisenberg_neighbours", n_shells_exchange=2, jij=[1.0, 0.5], ddi_method="none")
    p = _state(tmp_path, cfg)
    try:
        idx, tr, J = hamiltonian.get_exchange_pairs(p)
        S = oracle.System(cfg)
        want = [list(S.params.exch[k]) for k in range(S.params.n_exch)]
        got = [[int(a), int(b)] + [int(x) for x in t] for (a, b), t in zip(idx, tr)]
        assert got == want
        np.testing.assert_array_equal(J, np.array([S.params.exch_J[k] for k in range(S.params.n_exch)], np.float32))
    finally:
        state.delete(p)


def test_host_parameters_round_trip(lib, tmp_path):
    from spirit_amd import geometry, hamiltonian, state
    from spirit_amd.parameters import llg
    from spirit_amd.configs import SQUARE_64
    p = _state(tmp_path, SQUARE_64)
    try:
        assert geometry.get_n_cells(p) == [64, 64, 1]
        assert geometry.get_nos(p) == 4096
        m, n = hamiltonian.get_field(p)
        assert abs(m - 2.0) < 1e-6 and n == [0.0, 0.0, 1.0]
        assert hamiltonian.get_boundary_conditions(p) == [True, True, False]
        assert abs(llg.get_timestep(p) - 1e-3) < 1e-9
        assert abs(llg.get_damping(p) - 0.3) < 1e-7
        hamiltonian.set_field(p, 1.5, [0, 1, 1])
        m, n = hamiltonian.get_field(p)
        assert abs(m - 1.5) < 1e-6 and abs(n[1] - 2 ** -0.5) < 1e-6
    finally:
        state.delete(p)


def test_api_errors_are_logged_not_raised(lib, tmp_path):
    from spirit_amd import log, simulation, state, system
    from spirit_amd.configs import SQUARE_64
    p = _state(tmp_path, SQUARE_64)
    try:
        n0 = log.get_n_warnings(p)
        assert system.get_energy(p, idx_image=5) == 0.0  # invalid index -> warning, default value
        assert log.get_n_warnings(p) == n0 + 1
        simulation.single_shot(p)  # no single-shot method started -> warning
        assert log.get_n_warnings(p) == n0 + 2
    finally:
        state.delete(p)
