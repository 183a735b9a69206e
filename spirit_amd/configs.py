"""Input files in the reference's keyword format, and the benchmark workloads.

``write_config(cfg, path)`` turns a dict of the reference's input keywords
(core/docs/Input.md; parsed by core/src/io/Configparser.cpp) into an input file
that ``state.setup`` reads.  Keys: bravais_lattice | bravais_vectors, basis,
n_basis_cells, lattice_constant, boundary_conditions, mu_s,
external_field_magnitude/normal, anisotropy_magnitude/normal,
cubic_anisotropy_magnitude, hamiltonian, n_shells_exchange, jij, n_shells_dmi,
dij, dm_chirality, pairs (heisenberg_pairs rows i j da db dc Dx Dy Dz J),
ddi_method, ddi_n_periodic_images, ddi_pb_zero_padding, ddi_radius, quadruplets (rows
i j da_j db_j dc_j k da_k db_k dc_k l da_l db_l dc_l Q), llg_* (including llg_beta,
llg_stt_magnitude / _use_gradient / _polarisation_normal and
llg_temperature_gradient_inclination / _direction).
"""


def _vec(v):
    return " ".join(repr(float(x)) if isinstance(x, float) else str(x) for x in v)


def config_text(cfg):
    lines = ["log_console_level %d" % int(cfg.get("log_console_level", 2))]
    if "bravais_vectors" in cfg:
        lines.append("bravais_vectors")
        for row in cfg["bravais_vectors"]:
            lines.append(_vec(row))
    else:
        lines.append("bravais_lattice %s" % cfg.get("bravais_lattice", "sc"))
    if "basis" in cfg:
        lines.append("basis")
        lines.append(str(len(cfg["basis"])))
        for row in cfg["basis"]:
            lines.append(_vec(row))
    lines.append("n_basis_cells %s" % _vec(cfg.get("n_basis_cells", [100, 100, 1])))
    if "lattice_constant" in cfg:
        lines.append("lattice_constant %r" % float(cfg["lattice_constant"]))
    mu = cfg.get("mu_s", [1.0])
    if not isinstance(mu, (list, tuple)):
        mu = [mu]
    lines.append("mu_s %s" % _vec([float(x) for x in mu]))
    lines.append("boundary_conditions %s" % _vec(cfg.get("boundary_conditions", [0, 0, 0])))
    lines.append("external_field_magnitude %r" % float(cfg.get("external_field_magnitude", 0.0)))
    lines.append("external_field_normal %s" % _vec([float(x) for x in cfg.get("external_field_normal", [0, 0, 1])]))
    lines.append("anisotropy_magnitude %r" % float(cfg.get("anisotropy_magnitude", 0.0)))
    lines.append("anisotropy_normal %s" % _vec([float(x) for x in cfg.get("anisotropy_normal", [0, 0, 1])]))
    if "cubic_anisotropy_magnitude" in cfg:
        lines.append("cubic_anisotropy_magnitude %r" % float(cfg["cubic_anisotropy_magnitude"]))
    ham = cfg.get("hamiltonian", "heisenberg_neighbours")
    lines.append("hamiltonian %s" % ham)
    if ham == "heisenberg_pairs":
        pairs = cfg.get("pairs", [])
        lines.append("n_interaction_pairs %d" % len(pairs))
        lines.append("i j da db dc Dijx Dijy Dijz Jij")
        for p in pairs:
            lines.append(" ".join(str(x) for x in p))
    else:
        nse = int(cfg.get("n_shells_exchange", 0))
        lines.append("n_shells_exchange %d" % nse)
        if nse:
            lines.append("jij %s" % _vec([float(x) for x in cfg["jij"]]))
        nsd = int(cfg.get("n_shells_dmi", 0))
        lines.append("n_shells_dmi %d" % nsd)
        if nsd:
            lines.append("dij %s" % _vec([float(x) for x in cfg["dij"]]))
        lines.append("dm_chirality %d" % int(cfg.get("dm_chirality", 1)))
    lines.append("ddi_method %s" % cfg.get("ddi_method", "none"))
    lines.append("ddi_n_periodic_images %s" % _vec(cfg.get("ddi_n_periodic_images", [4, 4, 4])))
    lines.append("ddi_pb_zero_padding %d" % int(cfg.get("ddi_pb_zero_padding", 1)))
    lines.append("ddi_radius %r" % float(cfg.get("ddi_radius", 0.0)))
    if cfg.get("quadruplets"):
        quads = cfg["quadruplets"]
        lines.append("n_interaction_quadruplets %d" % len(quads))
        lines.append("i j da_j db_j dc_j k da_k db_k dc_k l da_l db_l dc_l Q")
        for q in quads:
            lines.append(" ".join(repr(float(x)) if isinstance(x, float) else str(x) for x in q))
    for k in ("llg_seed", "llg_dt", "llg_damping", "llg_temperature", "llg_force_convergence", "llg_n_iterations",
              "llg_n_iterations_log", "llg_n_iterations_amortize", "gneb_spring_constant", "gneb_force_convergence",
              "gneb_n_iterations", "gneb_n_iterations_log", "gneb_moving_endpoints", "gneb_translating_endpoints",
              "gneb_equilibrium_delta_Rx_left", "gneb_equilibrium_delta_Rx_right", "gneb_escape_first",
              "gneb_n_energy_interpolations"):
        if k in cfg:
            lines.append("%s %s" % (k, repr(cfg[k]) if isinstance(cfg[k], float) else cfg[k]))
    # after llg_temperature: the reader takes the first line that starts with a keyword
    for k in ("llg_temperature_gradient_direction", "llg_stt_polarisation_normal"):
        if k in cfg:
            lines.append("%s %s" % (k, _vec([float(x) for x in cfg[k]])))
    for k in ("llg_beta", "llg_stt_magnitude", "llg_temperature_gradient_inclination"):
        if k in cfg:
            lines.append("%s %r" % (k, float(cfg[k])))
    if "llg_stt_use_gradient" in cfg:
        lines.append("llg_stt_use_gradient %d" % int(bool(cfg["llg_stt_use_gradient"])))
    return "\n".join(lines) + "\n"


def write_config(cfg, path):
    with open(path, "w") as f:
        f.write(config_text(cfg))
    return path


# ---------------------------------------------------------------------------
# Benchmark workloads (the reference's physics, one basis atom, mu_s = 2)
# ---------------------------------------------------------------------------
# Plain LLG, 64x64x1 square, periodic a,b: J=10, Neel DMI 6, K=0.2 z, B=2 T z, Depondt,
# dt=1e-3 ps, alpha=0.3, T=0.
SQUARE_64 = dict(bravais_lattice="sc", n_basis_cells=[64, 64, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
                 external_field_magnitude=2.0, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.2,
                 anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[10.0], n_shells_dmi=1, dij=[6.0],
                 dm_chirality=2, ddi_method="none", llg_dt=1e-3, llg_damping=0.3, llg_temperature=0.0,
                 llg_seed=20006, llg_force_convergence=1e-12)

# Plain LLG with dipolar FFT: 128x128x32 film, open, padded 256x256x64: J=5, Bloch DMI 1,
# K=0.05 z, B=0.5 T z, Heun, dt=5e-4 ps, alpha=0.1, T=10 K.
FILM_128 = dict(bravais_lattice="sc", n_basis_cells=[128, 128, 32], boundary_conditions=[0, 0, 0], mu_s=[2.0],
                external_field_magnitude=0.5, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.05,
                anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[5.0], n_shells_dmi=1, dij=[1.0],
                dm_chirality=1, ddi_method="fft", ddi_n_periodic_images=[0, 0, 0], ddi_pb_zero_padding=1,
                llg_dt=5e-4, llg_damping=0.1, llg_temperature=10.0, llg_seed=20006, llg_force_convergence=1e-12)

# Depondt + DDI on 256x256x128 (padded 512x512x256): J=5, Bloch DMI 1, per-site K_i in
# [0.02, 0.1] meV along z, B=0.2 T z, dt=5e-4 ps, alpha=0.3, T=0.
CUBE_256 = dict(bravais_lattice="sc", n_basis_cells=[256, 256, 128], boundary_conditions=[0, 0, 0], mu_s=[2.0],
                external_field_magnitude=0.2, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.0,
                n_shells_exchange=1, jij=[5.0], n_shells_dmi=1, dij=[1.0], dm_chirality=1, ddi_method="fft",
                ddi_n_periodic_images=[0, 0, 0], ddi_pb_zero_padding=1, llg_dt=5e-4, llg_damping=0.3,
                llg_temperature=0.0, llg_seed=20006, llg_force_convergence=1e-12)


def site_anisotropy(nos, seed, lo=0.02, hi=0.1):
    """Per-site K_i ~ U[lo, hi] meV from a seed."""
    import numpy as np
    return np.random.default_rng(seed).uniform(lo, hi, nos)


def random_spins(nos, seed):
    """i.i.d. uniform unit vectors from a seed (the workloads' initial state)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, nos)
    phi = rng.uniform(-np.pi, np.pi, nos)
    r = np.sqrt(1.0 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def scaled(cfg, n_cells, **kw):
    d = dict(cfg)
    d["n_basis_cells"] = list(n_cells)
    d.update(kw)
    return d


# Uneven ensemble: independent stochastic replicas on periodic L x L x 1 square lattices, L from
# {32, 48, 64, 96, 128, 192, 256}; J=10, Neel DMI 3, K=0.1 z, B ~ U[0, 2] T z, T ~ U[1, 50] K,
# alpha=0.2, Depondt, dt=1e-3 ps.
ENSEMBLE_SIDES = (32, 48, 64, 96, 128, 192, 256)


def ensemble(n_replicas=256, seed=20006):
    """The replica configurations of the uneven ensemble, drawn from a seed (replica k's thermal
    noise is seeded with seed + k)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = rng.choice(ENSEMBLE_SIDES, n_replicas)
    fields = rng.uniform(0.0, 2.0, n_replicas)
    temps = rng.uniform(1.0, 50.0, n_replicas)
    return [dict(bravais_lattice="sc", n_basis_cells=[int(L), int(L), 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
                 external_field_magnitude=float(B), external_field_normal=[0, 0, 1], anisotropy_magnitude=0.1,
                 anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[10.0], n_shells_dmi=1, dij=[3.0],
                 dm_chirality=2, ddi_method="none", llg_dt=1e-3, llg_damping=0.2, llg_temperature=float(T),
                 llg_seed=seed + k, llg_force_convergence=1e-12)
            for k, (L, B, T) in enumerate(zip(sides, fields, temps))]


# GNEB chain: 128x128 hexagonal (60 degrees) monolayer, periodic: J=10, Neel DMI 6, K=0.4 z,
# B=3 T z, mu_s=2; 32 images from an isolated Neel skyrmion (radius 8) to the ferromagnet,
# climbing image, VP.
HEX_128 = dict(bravais_lattice="hex2d", n_basis_cells=[128, 128, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=3.0, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.4,
               anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[10.0], n_shells_dmi=1, dij=[6.0],
               dm_chirality=2, ddi_method="none", llg_dt=1e-3, gneb_spring_constant=1.0,
               gneb_force_convergence=1e-12)
