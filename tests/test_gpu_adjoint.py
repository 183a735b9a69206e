"""The discrete adjoint of the Depondt / Heun runs (Spirit_HIP_Adjoint_*, spirit_amd.adjoint).

* Dot-product test: for the same run, the adjoint gradient paired with a random direction
  equals the tangent-linear run's directional derivative (both are exact derivatives of the
  same discrete map; agreement to 1e-9 relative).
* Central finite differences of the oracle's CPU forward run (oracle/spirit_oracle.c) on the
  32x32x8 dipolar sub-case of the 256x256x128 workload: dL/dJ, dL/dD, dL/dB and dL/dK_i of a
  mean-square loss to a seeded target texture agree to 1e-6 relative.
* The torch.autograd.Function returns the same gradients as the session's backward pass."""
import numpy as np
import pytest

from spirit_amd import configs

pytestmark = pytest.mark.gpu


def print_case(*a):
    print("CASE", *a, flush=True)


@pytest.fixture
def state(tmp_path):
    from spirit_amd import state as st
    made = []

    def make(cfg, site_K=None):
        from spirit_amd import hip
        path = str(tmp_path / "input.cfg")
        configs.write_config(cfg, path)
        p = st.setup(path, True)
        assert p
        if site_K is not None:
            hip.set_site_anisotropy(p, site_K, (0.0, 0.0, 1.0))
        made.append(p)
        return p

    yield make
    for p in made:
        st.delete(p)


FILM = dict(bravais_lattice="sc", n_basis_cells=[10, 8, 3], boundary_conditions=[0, 0, 0], mu_s=[2.0],
            external_field_magnitude=0.4, external_field_normal=[0.1, 0.0, 1.0], anisotropy_magnitude=0.05,
            anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[5.0], n_shells_dmi=1, dij=[1.0], dm_chirality=1,
            ddi_method="fft", ddi_n_periodic_images=[0, 0, 0], ddi_pb_zero_padding=1, llg_dt=5e-4,
            llg_damping=0.3, llg_temperature=0.0, llg_seed=7)


@pytest.mark.parametrize("solver,temperature", [(2, 0.0), (3, 0.0), (2, 20.0), (3, 20.0)])
def test_adjoint_matches_tangent_linear(state, solver, temperature):
    from spirit_amd import adjoint
    print_case("adjoint dot product", solver, temperature)
    cfg = dict(FILM, llg_temperature=temperature)
    n = 10 * 8 * 3
    rng = np.random.default_rng(5)
    K = rng.uniform(0.02, 0.1, n)
    p = state(cfg, site_K=K)
    sess = adjoint.Session(p, solver=solver, n_steps=30, checkpoint_every=7)
    assert (sess.n_exchange, sess.n_dmi, sess.n_aniso, sess.nos) == (1, 1, 1, n)
    s0 = configs.random_spins(n, 3)
    w = rng.normal(size=(n, 3))
    ds0 = rng.normal(size=(n, 3))
    dJ, dD, dB, dK = rng.normal(size=1), rng.normal(size=1), rng.normal(size=1), rng.normal(size=1)
    dKs = rng.normal(size=n)
    sN = sess.forward(s0)
    g = sess.backward(w)
    sN2, dsN = sess.tangent(s0, ds0, dJ, dD, dB, dK, dKs)
    np.testing.assert_allclose(sN2, sN, atol=1e-12)
    lhs = float(np.sum(w * dsN))
    rhs = float(np.sum(g["s0"] * ds0) + g["J"] @ dJ + g["D"] @ dD + g["B"] @ dB + g["K"] @ dK + g["K_site"] @ dKs)
    assert abs(lhs - rhs) <= 1e-9 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    # each parameter class separately (a wrong sign in one term cannot hide behind the others)
    for name, kw in [("J", dict(dJ=dJ)), ("D", dict(dD=dD)), ("B", dict(dB=dB)), ("K", dict(dK=dK)),
                     ("K_site", dict(dK_site=dKs))]:
        _, d = sess.tangent(s0, np.zeros((n, 3)), **kw)
        direction = list(kw.values())[0]
        a, b = float(np.sum(w * d)), float(g[name] @ direction)
        assert abs(a - b) <= 1e-9 * max(abs(a), abs(b), 1e-30), (name, a, b)


def test_adjoint_matches_oracle_finite_differences(state, oracle):
    """32x32x8, open, dipolar FFT (padded 64x64x16), per-site K_i ~ U[0.02, 0.1], Depondt,
    dt 5e-4 ps, alpha 0.3, T = 0, 40 steps, checkpoints every 10; L = mean |s_N - t|^2."""
    from spirit_amd import adjoint
    print_case("adjoint vs oracle finite differences", "32x32x8")
    cfg = dict(configs.CUBE_256, n_basis_cells=[32, 32, 8])
    n = 32 * 32 * 8
    K = configs.site_anisotropy(n, 9)
    steps = 40
    p = state(cfg, site_K=K)
    sess = adjoint.Session(p, solver=2, n_steps=steps, checkpoint_every=10)
    s0 = configs.random_spins(n, 21)
    target = configs.random_spins(n, 22)
    sN = sess.forward(s0)
    g = sess.backward(2.0 * (sN - target) / n)

    def run(cfg_, K_):
        S = oracle.System(cfg_, site_K=K_)
        s, _, _, _ = S.llg(s0, 2, steps, dt=cfg_["llg_dt"], damping=cfg_["llg_damping"])
        S.close()
        return s

    # the oracle run is the GPU run (sanity of the comparison itself)
    assert np.abs(run(cfg, K) - sN).max() < 1e-10

    def central(key, value, make):
        # five-point stencil, h = 1e-3 relative (truncation ~h^4); each L(+h) - L(-h) is formed
        # as mean((s+ - s-) . (s+ + s- - 2t)) from the state difference, so that a single-site
        # K_i (a 1e-6 effect on L ~ 2) is not lost to cancellation
        h = 1e-3 * max(abs(value), 0.01)

        def diff(e):
            sp, sm = run(*make(value + e)), run(*make(value - e))
            return float(np.mean(np.sum((sp - sm) * (sp + sm - 2 * target), axis=1)))
        return (8.0 * diff(h) - diff(2 * h)) / (12.0 * h)

    checks = [
        ("J", g["J"][0], central("J", cfg["jij"][0], lambda v: (dict(cfg, jij=[v]), K))),
        ("D", g["D"][0], central("D", cfg["dij"][0], lambda v: (dict(cfg, dij=[v]), K))),
        ("B", g["B"][0], central("B", cfg["external_field_magnitude"],
                                 lambda v: (dict(cfg, external_field_magnitude=v), K))),
    ]
    for i in (0, 4321, n - 1):
        def mk(v, i=i):
            K2 = K.copy()
            K2[i] = v
            return cfg, K2
        checks.append(("K_%d" % i, g["K_site"][i], central("K", K[i], mk)))
    for name, adj, fd in checks:
        print(name, adj, fd, flush=True)
        assert abs(adj - fd) <= 1e-6 * abs(fd), (name, adj, fd)


def test_torch_autograd_function(state):
    import torch
    from spirit_amd import adjoint
    print_case("adjoint", "torch autograd")
    n = 10 * 8 * 3
    p = state(FILM, site_K=np.full(n, 0.03))
    sess = adjoint.Session(p, solver=3, n_steps=12, checkpoint_every=5)
    dev = torch.device("cuda")
    s0 = torch.tensor(configs.random_spins(n, 1), dtype=torch.float64, device=dev, requires_grad=True)
    t = torch.tensor(configs.random_spins(n, 2), dtype=torch.float64, device=dev)
    J, D, B, Ks = sess.parameter_tensors(requires_grad=True, device=dev)
    sN = adjoint.LLGFunction.apply(sess, s0, J, D, B, Ks)
    assert sN.is_cuda and sN.shape == (n, 3)
    loss = ((sN - t) ** 2).sum(dim=1).mean()
    loss.backward()
    ref = sess.backward((2.0 * (sN.detach() - t) / n).cpu().numpy())
    np.testing.assert_allclose(s0.grad.cpu().numpy(), ref["s0"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(J.grad.cpu().numpy(), ref["J"], rtol=1e-12)
    np.testing.assert_allclose(D.grad.cpu().numpy(), ref["D"], rtol=1e-12)
    np.testing.assert_allclose(B.grad.cpu().numpy(), ref["B"], rtol=1e-12)
    np.testing.assert_allclose(Ks.grad.cpu().numpy(), ref["K_site"], rtol=1e-12, atol=1e-18)


def test_torch_function_applied_twice(state):
    """Two applications of LLGFunction on one session before a single backward (a loss that
    sums two runs from different initial states and parameters): each node differentiates
    its own run, so the gradients equal the sums of two fresh single-application runs."""
    import torch
    from spirit_amd import adjoint
    print_case("adjoint", "torch function applied twice")
    n = 10 * 8 * 3
    p = state(FILM, site_K=np.full(n, 0.03))
    sess = adjoint.Session(p, solver=2, n_steps=10, checkpoint_every=4)
    dev = torch.device("cuda")
    t = torch.tensor(configs.random_spins(n, 9), dtype=torch.float64, device=dev)
    base = sess.parameter_tensors(requires_grad=False, device=dev)

    def run(s0, scale):
        J, D, B, Ks = (x.clone().mul_(scale).requires_grad_(True) for x in base)
        return J, D, B, Ks, adjoint.LLGFunction.apply(sess, s0, J, D, B, Ks)

    starts = [torch.tensor(configs.random_spins(n, 20 + k), dtype=torch.float64, device=dev, requires_grad=True)
              for k in range(2)]
    scales = (1.0, 1.3)
    # both runs first, then one backward through their sum
    outs = [run(s0, c) for s0, c in zip(starts, scales)]
    loss = sum(((o[4] - t) ** 2).sum(dim=1).mean() for o in outs)
    loss.backward()
    for k, (s0, c) in enumerate(zip(starts, scales)):
        ref_s0 = s0.detach().clone().requires_grad_(True)
        J, D, B, Ks, sN = run(ref_s0, c)
        ((sN - t) ** 2).sum(dim=1).mean().backward()
        np.testing.assert_allclose(s0.grad.cpu().numpy(), ref_s0.grad.cpu().numpy(), rtol=1e-12, atol=1e-15)
        for a, b in zip(outs[k][:4], (J, D, B, Ks)):
            np.testing.assert_allclose(a.grad.cpu().numpy(), b.grad.cpu().numpy(), rtol=1e-12, atol=1e-18)
