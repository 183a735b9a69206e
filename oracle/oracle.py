"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

``params_from_config`` restates the config-key semantics of the reference's
parser (core/src/io/Configparser.cpp:271-567 geometry, 1256-1630 Heisenberg
Hamiltonian) for the keys the workloads use, so that a test can build the
oracle and the product from the same description.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_BASIS = 8
MAX_PAIRS = 256

MU_B = 0.057883817555
MU_0 = 2.0133545 * 1e-28
K_B = 0.08617330350
GAMMA = 0.1760859644

SOLVER_VP, SOLVER_SIB, SOLVER_DEPONDT, SOLVER_HEUN, SOLVER_RK4 = 0, 1, 2, 3, 4
DDI_NONE, DDI_FFT, DDI_CUTOFF = 0, 1, 3


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int * 3),
        ("n_cell_atoms", ctypes.c_int),
        ("bravais", (ctypes.c_double * 3) * 3),
        ("cell_atoms", (ctypes.c_double * 3) * MAX_BASIS),
        ("lattice_constant", ctypes.c_double),
        ("mu_s", ctypes.c_double * MAX_BASIS),
        ("bc", ctypes.c_int * 3),
        ("field_tesla", ctypes.c_double),
        ("field_normal", ctypes.c_double * 3),
        ("n_aniso", ctypes.c_int),
        ("aniso_idx", ctypes.c_int * MAX_BASIS),
        ("aniso_K", ctypes.c_double * MAX_BASIS),
        ("aniso_normal", (ctypes.c_double * 3) * MAX_BASIS),
        ("n_cubic", ctypes.c_int),
        ("cubic_idx", ctypes.c_int * MAX_BASIS),
        ("cubic_K", ctypes.c_double * MAX_BASIS),
        ("site_aniso_K", ctypes.POINTER(ctypes.c_double)),
        ("site_aniso_normal", ctypes.c_double * 3),
        ("n_exch", ctypes.c_int),
        ("exch", (ctypes.c_int * 5) * MAX_PAIRS),
        ("exch_J", ctypes.c_double * MAX_PAIRS),
        ("n_dmi", ctypes.c_int),
        ("dmi", (ctypes.c_int * 5) * MAX_PAIRS),
        ("dmi_D", ctypes.c_double * MAX_PAIRS),
        ("dmi_normal", (ctypes.c_double * 3) * MAX_PAIRS),
        ("ddi_method", ctypes.c_int),
        ("ddi_n_periodic_images", ctypes.c_int * 3),
        ("ddi_pb_zero_padding", ctypes.c_int),
        ("ddi_radius", ctypes.c_double),
        ("n_quad", ctypes.c_int),
        ("quad", (ctypes.c_int * 13) * 64),
        ("quad_Q", ctypes.c_double * 64),
    ]


class OrcLLG(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("damping", ctypes.c_double),
        ("temperature", ctypes.c_double),
        ("direct_minimization", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("noise", ctypes.POINTER(ctypes.c_double)),
        ("beta", ctypes.c_double),
        ("stt_use_gradient", ctypes.c_int),
        ("stt_magnitude", ctypes.c_double),
        ("stt_normal", ctypes.c_double * 3),
        ("tgrad_inclination", ctypes.c_double),
        ("tgrad_direction", ctypes.c_double * 3),
    ]


class OrcGNEB(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("spring_constant", ctypes.c_double),
        ("spring_force_ratio", ctypes.c_double),
        ("path_shortening_constant", ctypes.c_double),
        ("moving_endpoints", ctypes.c_int),
        ("translating_endpoints", ctypes.c_int),
        ("escape_first", ctypes.c_int),
        ("delta_Rx_left", ctypes.c_double),
        ("delta_Rx_right", ctypes.c_double),
        ("n_E_interpolations", ctypes.c_int),
        ("force_convergence", ctypes.c_double),
    ]


_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.orc_shell_radii.argtypes = [P(OrcParams), i, P(d)]
        L.orc_neighbours_in_shells.argtypes = [P(OrcParams), i, i, P(i), i]
        L.orc_neighbours_in_shells.restype = i
        L.orc_dmi_normal.argtypes = [P(OrcParams), P(i), i, P(d)]
        L.orc_setup_shells.argtypes = [P(OrcParams), i, P(d), i, P(d), i]
        L.orc_setup_shells.restype = i
        L.orc_make_pairs_redundant.argtypes = [P(OrcParams)]
        L.orc_idx_from_pair.argtypes = [i, P(i), P(i), i, P(i)]
        L.orc_idx_from_pair.restype = i
        L.orc_create.argtypes = [P(OrcParams)]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_nos.argtypes = [vp]
        L.orc_nos.restype = i
        L.orc_positions.argtypes = [vp, P(d)]
        L.orc_ddi_padded.argtypes = [vp, P(i)]
        L.orc_ddi_dipole_matrices.argtypes = [vp, P(d)]
        L.orc_gradient_and_energy.argtypes = [vp, P(d), P(d), P(d)]
        L.orc_gradient.argtypes = [vp, P(d), P(d)]
        L.orc_gradient_term.argtypes = [vp, i, P(d), P(d)]
        L.orc_hessian.argtypes = [vp, P(d), P(d)]
        L.orc_ddi_cutoff_pairs.argtypes = [vp, P(i), P(d), P(d), i]
        L.orc_ddi_cutoff_pairs.restype = i
        L.orc_jacobian_direction.argtypes = [vp, P(d), P(d), P(d)]
        L.orc_temperature_distribution.argtypes = [vp, P(d), d, d, P(d)]
        L.orc_llg_iterate.argtypes = [vp, i, P(OrcLLG), P(d), i, ctypes.c_uint64, P(d), P(d), P(d)]
        L.orc_llg_run.argtypes = [vp, i, P(OrcLLG), P(d), ctypes.c_long, d, P(d), P(d), P(d)]
        L.orc_llg_run.restype = ctypes.c_long
        L.orc_normal3.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P(d)]
        L.orc_fft.argtypes = [P(d), i, i]
        L.orc_set_fft_parallel.argtypes = [i]
        L.orc_rotate.argtypes = [P(d), P(d), d, P(d)]
        L.orc_sib_transform.argtypes = [P(d), P(d), P(d)]
        L.orc_gneb_run.argtypes = [vp, i, P(OrcGNEB), i, P(i), P(d), ctypes.c_long, i, P(d), P(d), P(d), P(d), P(d),
                                   P(d)]
        L.orc_gneb_run.restype = ctypes.c_long
        L.orc_cubic_hermite.argtypes = [P(d), P(d), P(d), i, i, P(d), P(d)]
        L.orc_homogeneous_rotation.argtypes = [P(d), i, i, i]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


# ---------------------------------------------------------------------------
# config-dict -> oracle parameters
# ---------------------------------------------------------------------------
BRAVAIS = {
    # core/src/data/Geometry.cpp:446-488
    "sc": [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
    "fcc": [[0.5, 0.0, 0.5], [0.5, 0.5, 0.0], [0.0, 0.5, 0.5]],
    "bcc": [[0.5, 0.5, -0.5], [-0.5, 0.5, -0.5], [0.5, -0.5, -0.5]],
    "hex2d": [[0.5 * math.sqrt(3), -0.5, 0.0], [0.5 * math.sqrt(3), 0.5, 0.0], [0.0, 0.0, 1.0]],
    "hex2d60": [[0.5 * math.sqrt(3), -0.5, 0.0], [0.5 * math.sqrt(3), 0.5, 0.0], [0.0, 0.0, 1.0]],
    "hex2d120": [[0.5, -0.5 * math.sqrt(3), 0.0], [0.5, 0.5 * math.sqrt(3), 0.0], [0.0, 0.0, 1.0]],
}


def _normalized(v):
    v = [float(x) for x in v]
    z = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    if z > 0:
        n = math.sqrt(z)
        v = [x / n for x in v]
    return v


def params_from_config(cfg, site_K=None):
    """Build OrcParams from a config dict (same keys as the reference's input file).

    Returns (params, keepalive) -- keepalive holds arrays referenced by pointer.
    """
    p = OrcParams()
    bv = cfg.get("bravais_vectors")
    if bv is None:
        bv = BRAVAIS[cfg.get("bravais_lattice", "sc").lower()]
    for k in range(3):
        for c in range(3):
            p.bravais[k][c] = float(bv[k][c])
    basis = cfg.get("basis", [[0.0, 0.0, 0.0]])
    p.n_cell_atoms = len(basis)
    for a, pos in enumerate(basis):
        for c in range(3):
            p.cell_atoms[a][c] = float(pos[c])
    nc = cfg.get("n_basis_cells", [100, 100, 1])
    for c in range(3):
        p.n_cells[c] = int(nc[c])
    p.lattice_constant = float(cfg.get("lattice_constant", 1.0))
    mu = cfg.get("mu_s", [1.0])
    if not isinstance(mu, (list, tuple)):
        mu = [mu]
    for a in range(p.n_cell_atoms):
        p.mu_s[a] = float(mu[a] if a < len(mu) else mu[0])
    bc = cfg.get("boundary_conditions", [0, 0, 0])
    for c in range(3):
        p.bc[c] = 1 if int(bc[c]) != 0 else 0

    p.field_tesla = float(cfg.get("external_field_magnitude", 0.0))
    bn = _normalized(cfg.get("external_field_normal", [0, 0, 1]))
    if math.sqrt(sum(x * x for x in bn)) < 1e-8:
        bn = [0.0, 0.0, 1.0]
    for c in range(3):
        p.field_normal[c] = bn[c]

    K = float(cfg.get("anisotropy_magnitude", 0.0))
    kn = _normalized(cfg.get("anisotropy_normal", [0, 0, 1]))
    if K != 0:
        p.n_aniso = p.n_cell_atoms
        for a in range(p.n_cell_atoms):
            p.aniso_idx[a] = a
            p.aniso_K[a] = K
            for c in range(3):
                p.aniso_normal[a][c] = kn[c]
    K4 = float(cfg.get("cubic_anisotropy_magnitude", 0.0))
    if K4 != 0:
        p.n_cubic = p.n_cell_atoms
        for a in range(p.n_cell_atoms):
            p.cubic_idx[a] = a
            p.cubic_K[a] = K4
    keep = []
    if site_K is not None:
        arr = np.ascontiguousarray(site_K, dtype=np.float64)
        keep.append(arr)
        p.site_aniso_K = _dp(arr)
        sn = _normalized(cfg.get("site_anisotropy_normal", [0, 0, 1]))
        for c in range(3):
            p.site_aniso_normal[c] = sn[c]

    ham = cfg.get("hamiltonian", "heisenberg_neighbours")
    if ham == "heisenberg_pairs":
        ne = nd = 0
        for row in cfg.get("pairs", []):
            i, j, da, db, dc, Dx, Dy, Dz, J = row
            dn = math.sqrt(Dx ** 2 + Dy ** 2 + Dz ** 2)
            if dn != 0:
                Dx, Dy, Dz = Dx / dn, Dy / dn, Dz / dn
            if J != 0:
                p.exch[ne][:] = [i, j, da, db, dc]
                p.exch_J[ne] = J
                ne += 1
            if dn != 0:
                p.dmi[nd][:] = [i, j, da, db, dc]
                p.dmi_D[nd] = dn
                p.dmi_normal[nd][:] = [Dx, Dy, Dz]
                nd += 1
        p.n_exch, p.n_dmi = ne, nd
        lib().orc_make_pairs_redundant(ctypes.byref(p))
    else:
        nse = int(cfg.get("n_shells_exchange", 0))
        nsd = int(cfg.get("n_shells_dmi", 0))
        jij = np.array(cfg.get("jij", [0.0] * max(nse, 1)), dtype=np.float64)
        dij = np.array(cfg.get("dij", [0.0] * max(nsd, 1)), dtype=np.float64)
        rc = lib().orc_setup_shells(ctypes.byref(p), nse, _dp(jij), nsd, _dp(dij), int(cfg.get("dm_chirality", 1)))
        assert rc == 0

    m = cfg.get("ddi_method", "none")
    p.ddi_method = {"none": DDI_NONE, "fft": DDI_FFT, "cutoff": DDI_CUTOFF}[m]
    im = cfg.get("ddi_n_periodic_images", [4, 4, 4])
    for c in range(3):
        p.ddi_n_periodic_images[c] = int(im[c])
    p.ddi_pb_zero_padding = 1 if int(cfg.get("ddi_pb_zero_padding", 1)) else 0
    p.ddi_radius = float(cfg.get("ddi_radius", 0.0))
    nq = 0
    for row in cfg.get("quadruplets", []):
        if float(row[13]) == 0:
            continue
        r = [int(x) for x in row[:13]]  # i j da_j db_j dc_j k da_k db_k dc_k l da_l db_l dc_l
        p.quad[nq][:] = [r[0], r[1], r[5], r[9]] + r[2:5] + r[6:9] + r[10:13]
        p.quad_Q[nq] = float(row[13])
        nq += 1
    p.n_quad = nq
    return p, keep


class System:
    """An oracle spin system (one image)."""

    def __init__(self, cfg, site_K=None):
        self.params, self._keep = params_from_config(cfg, site_K)
        self.cfg = cfg
        self.h = lib().orc_create(ctypes.byref(self.params))
        self.nos = lib().orc_nos(self.h)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def positions(self):
        out = np.zeros((self.nos, 3))
        lib().orc_positions(self.h, _dp(out))
        return out

    def padded(self):
        out = np.zeros(3, dtype=np.int32)
        lib().orc_ddi_padded(self.h, _ip(out))
        return out

    def gradient_and_energy(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(self.nos, 3)
        g = np.zeros((self.nos, 3))
        e = ctypes.c_double(0)
        lib().orc_gradient_and_energy(self.h, _dp(s), _dp(g), ctypes.byref(e))
        return g, e.value

    def gradient(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(self.nos, 3)
        g = np.zeros((self.nos, 3))
        lib().orc_gradient(self.h, _dp(s), _dp(g))
        return g

    def gradient_term(self, term, spins):
        s = np.ascontiguousarray(spins, dtype=np.float64).reshape(self.nos, 3)
        g = np.zeros((self.nos, 3))
        lib().orc_gradient_term(self.h, int(term), _dp(s), _dp(g))
        return g

    def hessian(self, spins=None):
        """Dense embedding Hessian (3N x 3N), Hamiltonian_Heisenberg::Hessian."""
        s = np.zeros((self.nos, 3)) if spins is None else np.ascontiguousarray(spins, dtype=np.float64)
        H = np.zeros((3 * self.nos, 3 * self.nos))
        lib().orc_hessian(self.h, _dp(s), _dp(H))
        return H

    def jacobian_direction(self, vf, je):
        """Vectormath::jacobian of the field vf (nos, 3) applied to the direction je."""
        v = np.ascontiguousarray(vf, dtype=np.float64).reshape(self.nos, 3)
        out = np.zeros((self.nos, 3))
        lib().orc_jacobian_direction(self.h, _dp(v), _dp(np.asarray(je, dtype=np.float64)), _dp(out))
        return out

    def temperature_distribution(self, direction, T, inclination):
        """Per-site temperature of a gradient (Vectormath::get_gradient_distribution)."""
        out = np.zeros(self.nos)
        lib().orc_temperature_distribution(self.h, _dp(np.asarray(direction, dtype=np.float64)), ctypes.c_double(T),
                                           ctypes.c_double(inclination), _dp(out))
        return out

    def llg(self, spins, solver, n_steps, dt=1e-3, damping=0.3, temperature=0.0, seed=0, step0=0,
            direct_minimization=False, noise=None, beta=0.0, stt=None, temperature_gradient=None):
        """stt = (magnitude, normal, use_gradient); temperature_gradient = (inclination, direction)."""
        s = np.array(spins, dtype=np.float64).reshape(self.nos, 3).copy()
        par = OrcLLG()
        par.dt, par.damping, par.temperature = dt, damping, temperature
        par.direct_minimization = 1 if direct_minimization else 0
        par.seed = seed
        par.beta = beta
        if stt is not None:
            mag, normal, grad = stt
            par.stt_magnitude = mag
            par.stt_normal[:] = _normalized(normal)
            par.stt_use_gradient = 1 if grad else 0
        if temperature_gradient is not None:
            par.tgrad_inclination = temperature_gradient[0]
            par.tgrad_direction[:] = [float(x) for x in temperature_gradient[1]]
        keep = None
        if noise is not None:
            keep = np.ascontiguousarray(noise, dtype=np.float64)
            par.noise = _dp(keep)
        e = ctypes.c_double(0)
        t = ctypes.c_double(0)
        eff = np.zeros((self.nos, 3))
        lib().orc_llg_iterate(self.h, int(solver), ctypes.byref(par), _dp(s), int(n_steps), int(step0),
                              ctypes.byref(e), ctypes.byref(t), _dp(eff))
        del keep
        return s, e.value, t.value, eff

    def llg_run(self, spins, solver, n_iterations, force_convergence, dt=1e-3, damping=0.3,
                direct_minimization=False):
        """Method::Iterate with the convergence test every iteration; returns (spins, energy,
        max torque, effective field, iterations done)."""
        s = np.array(spins, dtype=np.float64).reshape(self.nos, 3).copy()
        par = OrcLLG()
        par.dt, par.damping, par.temperature = dt, damping, 0.0
        par.direct_minimization = 1 if direct_minimization else 0
        e = ctypes.c_double(0)
        t = ctypes.c_double(0)
        eff = np.zeros((self.nos, 3))
        n = lib().orc_llg_run(self.h, int(solver), ctypes.byref(par), _dp(s), int(n_iterations),
                              float(force_convergence), ctypes.byref(e), ctypes.byref(t), _dp(eff))
        return s, e.value, t.value, eff, n

    def gneb(self, chain, solver, n_iterations, image_type=None, dt=1e-3, spring_constant=1.0,
             spring_force_ratio=0.0, path_shortening_constant=0.0, moving_endpoints=False,
             translating_endpoints=False, escape_first=False, delta_Rx=(1.0, 1.0), n_E_interpolations=10,
             force_convergence=1e-12, check_convergence=True):
        """Method_GNEB run on a (noi, nos, 3) chain; returns (chain, dict of chain data, iterations)."""
        c = np.array(chain, dtype=np.float64).reshape(-1, self.nos, 3).copy()
        noi = c.shape[0]
        it = np.zeros(noi, dtype=np.int32) if image_type is None else np.ascontiguousarray(image_type, np.int32)
        par = OrcGNEB()
        par.dt, par.spring_constant, par.spring_force_ratio = dt, spring_constant, spring_force_ratio
        par.path_shortening_constant = path_shortening_constant
        par.moving_endpoints, par.translating_endpoints = int(moving_endpoints), int(translating_endpoints)
        par.escape_first = int(escape_first)
        par.delta_Rx_left, par.delta_Rx_right = delta_Rx
        par.n_E_interpolations, par.force_convergence = n_E_interpolations, force_convergence
        E, Rx, mt = np.zeros(noi), np.zeros(noi), np.zeros(noi)
        eff = np.zeros((noi, self.nos, 3))
        npts = noi + (noi - 1) * n_E_interpolations
        Rxi, Ei = np.zeros(npts), np.zeros(npts)
        n = lib().orc_gneb_run(self.h, int(solver), ctypes.byref(par), noi, _ip(it), _dp(c), int(n_iterations),
                               1 if check_convergence else 0, _dp(E), _dp(Rx), _dp(mt), _dp(eff), _dp(Rxi), _dp(Ei))
        data = dict(energies=E, Rx=Rx, max_torque=mt, eff=eff, Rx_interpolated=Rxi, E_interpolated=Ei)
        return c, data, n


def cubic_hermite(x, p, m, n_interp):
    x, p, m = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, p, m))
    n = len(x)
    npts = n + (n - 1) * n_interp
    ox, op = np.zeros(npts), np.zeros(npts)
    lib().orc_cubic_hermite(_dp(x), _dp(p), _dp(m), n, n_interp, _dp(ox), _dp(op))
    return ox, op


def homogeneous_rotation(chain, idx_1, idx_2):
    c = np.array(chain, dtype=np.float64).copy()
    lib().orc_homogeneous_rotation(_dp(c), c.shape[1], idx_1, idx_2)
    return c


def neighbours_in_shells(cfg, n_shells, redundant=True):
    p, _ = params_from_config(dict(cfg, n_shells_exchange=0, n_shells_dmi=0))
    out = np.zeros((MAX_PAIRS, 6), dtype=np.int32)
    n = lib().orc_neighbours_in_shells(ctypes.byref(p), n_shells, 1 if redundant else 0, _ip(out), MAX_PAIRS)
    assert n >= 0
    return out[:n].copy()


def fft(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft(a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a), sign)
    return a


def set_fft_parallel(on):
    """True: the lines of each 3-D transform run on all OpenMP threads; False: each transform
    runs on one thread, as the reference's kissFFT build does."""
    lib().orc_set_fft_parallel(1 if on else 0)


def normal3(seed, step, site):
    out = np.zeros(3)
    lib().orc_normal3(seed, step, site, _dp(out))
    return out


def random_spins(nos, seed):
    """i.i.d. uniform unit vectors (numpy Generator), the workloads' injected initial state."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, nos)
    phi = rng.uniform(-math.pi, math.pi, nos)
    r = np.sqrt(1.0 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
