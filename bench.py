#!/usr/bin/env python3
"""Benchmark of the HIP spin-dynamics engine on one GPU.

Workload (default): Depondt LLG with dipole-dipole interaction by zero-padded FFT
convolution on a 256x256x128 simple-cubic lattice (open boundaries, padded grid
512x512x256), J=5 meV, Bloch DMI 1 meV, per-site uniaxial K_i ~ U[0.02, 0.1] meV
along z, B=0.2 T along z, mu_s=2, dt=5e-4 ps, alpha=0.3, T=0, FP64, random initial
spins from a seed.  A "step" is one full Depondt iteration (two force evaluations,
each with the FFT convolution) followed by the method's post-iteration hook (energy,
maximum torque and effective field of every step, as Simulation_N_Shot runs it) on the
spins already resident in HBM.

Prints one JSON line: spin-steps/s, the roofline of the dominant kernel (the DDI
convolution pass, timed live with HIP events on the engine's stream) and the CPU
oracle timed on the host cores on a bounded sample of the same workload.

Two more legs report into the same line: "batched" -- 256 independent stochastic
replicas on L x L x 1 square lattices (L from {32..256}, per-replica B and T,
Depondt, 500 steps) through one Spirit_HIP_Simulation_LLG_Start_Batch against the
same replicas run one Simulation_LLG_Start after another; "adjoint" -- 500 steps of
the default workload forward with checkpoints every 50 steps plus the discrete
adjoint of a mean-square loss (gradients for J, D, B and every K_i); "gneb" -- a
32-image chain on the 128x128 hexagonal monolayer (skyrmion collapse, climbing
image, VP), 2000 iterations; "minimiser" -- 200 LBFGS_OSO iterations of the default
workload; "mc" -- 1000 Metropolis sweeps on the default lattice without DDI.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "cube256_depondt_ddi": dict(cfg="CUBE_256", solver=2, site_K=True, dims=(256, 256, 128)),
    "film128_heun_ddi": dict(cfg="FILM_128", solver=3, site_K=False, dims=(128, 128, 32)),
    "square64_depondt": dict(cfg="SQUARE_64", solver=2, site_K=False, dims=(64, 64, 1)),
}


def alg_bytes_per_spin_step(dims, solver, ddi):
    """Algorithmic HBM bytes per spin-step (FP64, SoA): every distinct array read once and
    written once per stage; the DDI convolution as the zero-padding-aware 3-pass dataflow."""
    local = {2: 168.0, 3: 168.0, 1: 120.0, 4: 408.0}[solver]
    n_fe = 4 if solver == 4 else 2
    if not ddi:
        return local, 0.0
    Na, Nb, Nc = dims
    Pa, Pb, Pc = 2 * Na if Na > 1 else 1, 2 * Nb if Nb > 1 else 1, 2 * Nc if Nc > 1 else 1
    Ha = Pa // 2 + 1
    N = Na * Nb * Nc
    H_NN, H_PN, H_PP = Ha * Nb * Nc, Ha * Pb * Nc, Ha * Pb * Pc
    A = 24 * N + 48 * H_NN
    B = 48 * H_NN + 48 * H_PN
    C = 96 * H_PN + 96 * H_PP
    per_fe = (2 * A + 2 * B + C) / N
    return local + n_fe * (per_fe + 24.0), C


def design_bytes_per_spin_step(dims, solver, pass_c_launch_bytes):
    """HBM bytes per spin-step this engine's dataflow has to move (not the algorithmic
    figure, which prices the kernel spectrum as 6 dense complex components): the local part
    of the solver, the effective field the post-iteration hook writes on every step (24 B),
    and per force evaluation the DDI passes with pass C at its stored-spectrum bytes plus the
    stage's read of the field."""
    local = {2: 168.0, 3: 168.0, 1: 120.0, 4: 408.0}[solver] + 24.0
    n_fe = 4 if solver == 4 else 2
    Na, Nb, Nc = dims
    Pb = 2 * Nb if Nb > 1 else 1
    Ha = (2 * Na if Na > 1 else 1) // 2 + 1
    N = Na * Nb * Nc
    H_NN, H_PN = Ha * Nb * Nc, Ha * Pb * Nc
    per_fe = (2 * (24 * N + 48 * H_NN) + 2 * (48 * H_NN + 48 * H_PN) + pass_c_launch_bytes) / N
    return local + n_fe * (per_fe + 24.0)


def ddi_bytes_per_fe(dims):
    """Algorithmic bytes per spin of one DDI convolution (2A + 2B + C, without the stage's read
    of the field): pack + R2C along a, C2C along b, the c pass with the kernel multiply, and
    the inverse passes."""
    Na, Nb, Nc = dims
    Pa, Pb, Pc = 2 * Na if Na > 1 else 1, 2 * Nb if Nb > 1 else 1, 2 * Nc if Nc > 1 else 1
    Ha = Pa // 2 + 1
    N = Na * Nb * Nc
    H_NN, H_PN, H_PP = Ha * Nb * Nc, Ha * Pb * Nc, Ha * Pb * Pc
    return (2 * (24 * N + 48 * H_NN) + 2 * (48 * H_NN + 48 * H_PN) + 96 * H_PN + 96 * H_PP) / N


def adjoint_bytes_per_spin_step(dims, ckpt):
    """Algorithmic bytes per spin and step of the differentiable Depondt+DDI run (forward with
    checkpoints, then the backward pass), 24 B per vector field read or written:
    forward 216 + 2 DDI (the plain step) + 24/ckpt (checkpoint);
    recompute 312 + 2 DDI (the plain step's 9 field accesses plus the tape: F1, fv1, F2, fv2 and
    the step's spins) + 48/ckpt (the segment's checkpoint copied to the tape);
    adjoint 768 + 2 DDI (k_adj_predictor 3, k_adj_final 10, k_adj_predictor_back 9 and two
    k_adj_hessian 5 field accesses, and the two transposed convolutions)."""
    ddi = ddi_bytes_per_fe(dims)
    return 216 + 312 + 768 + 6 * ddi + 72.0 / ckpt


def run_gpu(args, wl):
    from spirit_amd import configs, hip, simulation, state, system
    cfg = dict(getattr(configs, wl["cfg"]))
    Na, Nb, Nc = wl["dims"]
    nos = Na * Nb * Nc
    path = os.path.join(tempfile.mkdtemp(), "bench.cfg")
    configs.write_config(cfg, path)
    p = state.setup(path, True)
    if not p:
        raise RuntimeError("State_Setup failed")
    if wl["site_K"]:
        hip.set_site_anisotropy(p, configs.site_anisotropy(nos, 20006), (0.0, 0.0, 1.0))
    system.get_spin_directions(p)[:] = configs.random_spins(nos, 20006)
    hip.set_host_sync(p, 0)
    t0 = time.perf_counter()
    simulation.start(p, simulation.METHOD_LLG, wl["solver"], single_shot=True)
    setup_s = time.perf_counter() - t0
    print("bench: setup %.2f s, device %s" % (setup_s, hip.device_name()), file=sys.stderr, flush=True)
    if args.warmup > 0:
        simulation.n_shot(p, args.warmup)
    hip.profile_enable(p, True)
    hip.profile_reset(p)
    t0 = time.perf_counter()
    simulation.n_shot(p, args.steps)  # returns after the stream has drained
    t1 = time.perf_counter()
    info = hip.ddi_info(p)
    conv_ms, conv_n = hip.profile_get(p, "ddi_conv")
    stage_ms, stage_n = hip.profile_get(p, "llg_stage")
    print("bench: %d steps in %.3f s" % (args.steps, t1 - t0), file=sys.stderr, flush=True)
    s = system.get_spin_directions(p)
    import numpy as np
    norm_err = float(np.abs(np.linalg.norm(s, axis=1) - 1).max())
    simulation.stop(p)
    state.delete(p)
    return dict(seconds=t1 - t0, nos=nos, conv_ms=conv_ms, conv_n=conv_n, stage_ms=stage_ms, stage_n=stage_n,
                norm_err=norm_err, setup_s=setup_s, ddi_info=info)


def copy_rates(sizes_mib=(256, 1024, 4096), reps=20):
    """Device bandwidth of plain streaming kernels (torch copy, sum, fill; FP64) with HIP
    events: the practical ceiling the DDI and stage kernels are compared with."""
    import torch
    dev = torch.device("cuda")
    out = {}
    for mib in sizes_mib:
        n = mib * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        res = {}
        for name, fn, moved in (("copy", lambda: b.copy_(a), 2), ("read", lambda: a.sum(), 1),
                                ("write", lambda: b.fill_(1.5), 1)):
            fn()
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res[name + "_GBps"] = moved * n * 8 / ms / 1e6
        out["%d_MiB" % mib] = res
        del a, b
    return out


def copy_calibration(mib):
    """Device-to-device copies of `mib` MiB: under a counter profiler, the copy kernel's
    FETCH_SIZE / WRITE_SIZE per dispatch against its known size give the counters' scale."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def run_batched(n_steps, n_replicas, sequential=True):
    """The uneven ensemble (n_replicas independent stochastic replicas, L x L x 1 with L from
    {32..256}, per-replica B and T, Depondt dt=1e-3 ps, n_steps steps): every replica run by its
    own Simulation_LLG_Start in turn, then all replicas by one Spirit_HIP_Simulation_LLG_Start_Batch.
    Both timings are the API calls end to end (device set-up, run, spins back in the states)."""
    from spirit_amd import configs, hip, simulation, state, system
    cfgs = configs.ensemble(n_replicas)
    tmp = tempfile.mkdtemp()
    states, spins0 = [], []
    t0 = time.perf_counter()
    for k, cfg in enumerate(cfgs):
        path = os.path.join(tmp, "r%d.cfg" % k)
        configs.write_config(cfg, path)
        p = state.setup(path, True)
        if not p:
            raise RuntimeError("State_Setup failed")
        states.append(p)
        spins0.append(configs.random_spins(system.get_nos(p), 30000 + k))
    print("bench: ensemble of %d replicas set up in %.1f s" % (n_replicas, time.perf_counter() - t0), file=sys.stderr,
          flush=True)
    nos = sum(len(s) for s in spins0)

    def reset():
        for p, s in zip(states, spins0):
            system.get_spin_directions(p)[:] = s

    # warm-up of both paths on a few steps (module loading, first allocations)
    reset()
    simulation.start(states[0], simulation.METHOD_LLG, 2, n_iterations=2)
    hip.llg_start_batch(states[:2], 2, n_iterations=2)
    t_seq = None
    if sequential:
        reset()
        t0 = time.perf_counter()
        for p in states:
            simulation.start(p, simulation.METHOD_LLG, 2, n_iterations=n_steps)
        t_seq = time.perf_counter() - t0
        print("bench: sequential ensemble %.3f s" % t_seq, file=sys.stderr, flush=True)
    reset()
    t0 = time.perf_counter()
    infos = hip.llg_start_batch(states, 2, n_iterations=n_steps)
    t_bat = time.perf_counter() - t0
    print("bench: batched ensemble %.3f s" % t_bat, file=sys.stderr, flush=True)
    done = all(i.total_iterations == n_steps for i in infos)
    for p in states:
        state.delete(p)
    value = nos * n_steps / t_bat
    ach = 168.0 * value / 1e9
    return {"workload": "ensemble%d_depondt" % n_replicas, "replicas": n_replicas, "spins": nos, "steps": n_steps,
            "value": value, "sequential_value": nos * n_steps / t_seq if t_seq else None,
            "unit": "spin-steps/s", "speedup_vs_sequential": t_seq / t_bat if t_seq else None, "batched_s": t_bat,
            "sequential_s": t_seq,
            "all_replicas_completed": done,
            "roofline": {"bound": "hbm", "alg_bytes_per_spin_step": 168.0, "achieved": ach, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                         "timed": "the whole Spirit_HIP_Simulation_LLG_Start_Batch call (set-up, copies, run)"}}


def run_adjoint(n_steps, ckpt, plain_ms_per_step):
    """The differentiable Depondt+DDI run on 256x256x128 (per-site K_i, B=0.2 T, dt 5e-4 ps,
    alpha 0.3, T=0): forward with a checkpoint every `ckpt` steps, then the discrete adjoint of
    L = mean |s_N - t|^2 to a seeded target t, spins and dL/ds on the GPU."""
    import torch
    from spirit_amd import adjoint, configs, hip, state
    cfg = dict(configs.CUBE_256)
    nos = 256 * 256 * 128
    path = os.path.join(tempfile.mkdtemp(), "adjoint.cfg")
    configs.write_config(cfg, path)
    p = state.setup(path, True)
    if not p:
        raise RuntimeError("State_Setup failed")
    hip.set_site_anisotropy(p, configs.site_anisotropy(nos, 20006), (0.0, 0.0, 1.0))
    t0 = time.perf_counter()
    sess = adjoint.Session(p, solver=2, n_steps=n_steps, checkpoint_every=ckpt)
    print("bench: adjoint session set up in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    dev = torch.device("cuda")
    s0 = torch.tensor(configs.random_spins(nos, 20006), device=dev)
    target = torch.tensor(configs.random_spins(nos, 20007), device=dev)
    times = []
    for rep in range(2):  # the first pass is the warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sN = sess.forward(s0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lam = 2.0 * (sN - target) / nos
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g = sess.backward(lam)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        times.append((t1 - t0, t3 - t2))
        print("bench: adjoint pass %d: forward %.3f s, backward %.3f s" % (rep, t1 - t0, t3 - t2), file=sys.stderr,
              flush=True)
    fwd, bwd = times[-1]
    loss = float(((sN - target) ** 2).sum(dim=1).mean())
    sess.close()
    state.delete(p)
    plain = plain_ms_per_step * 1e-3 * n_steps if plain_ms_per_step else None
    b_alg = adjoint_bytes_per_spin_step((256, 256, 128), ckpt)
    value = nos * n_steps / (fwd + bwd)
    ach = b_alg * value / 1e9
    return {"workload": "cube256_depondt_ddi_adjoint", "steps": n_steps, "checkpoint_every": ckpt, "spins": nos,
            "value": value, "unit": "spin-steps/s (forward + backward)",
            "roofline": {"bound": "hbm", "alg_bytes_per_spin_step": b_alg, "achieved": ach, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach / HBM_PEAK_GBS},
            "forward_s": fwd, "backward_s": bwd, "backward_over_forward": bwd / fwd,
            "gradient_over_plain_forward": (fwd + bwd) / plain if plain else None,
            "loss": loss, "dL_dJ": float(g["J"][0]), "dL_dD": float(g["D"][0]), "dL_dB": float(g["B"][0])}


# Algorithmic bytes of one GNEB VP iteration per spin (FP64, each array of the reference's
# dataflow read once and written once per pass; neighbouring images are the same arrays):
# gradient and distance (read s, write the effective field: 48), tangents (read s and the
# field, write tau and the gradient force: 96), total force and VP velocity (read tau, the
# gradient force, v and the previous force, write F_total and v: 144), VP update and hook
# (read F_total and s, write v, s and the projected force: 120).
GNEB_VP_BYTES_PER_SPIN_ITERATION = 408.0


def run_gneb(n_iter, noi):
    """GNEB on the 128x128 hexagonal monolayer: noi images from a Neel skyrmion (radius 8) to
    the ferromagnet by homogeneous rotation, climbing image set automatically after a first
    relaxation, VP; timed over n_iter iterations of one Simulation_GNEB_Start."""
    import numpy as np
    from spirit_amd import chain, configs, configuration, simulation, state, transition
    from spirit_amd.parameters import gneb
    path = os.path.join(tempfile.mkdtemp(), "gneb.cfg")
    configs.write_config(configs.HEX_128, path)
    p = state.setup(path, True)
    if not p:
        raise RuntimeError("State_Setup failed")
    configuration.plus_z(p)
    configuration.skyrmion(p, 8.0, 1.0, 0.0, False, False, False)
    chain.image_to_clipboard(p)
    for _ in range(noi - 1):
        chain.insert_image_after(p)
    chain.jump_to_image(p, noi - 1)
    configuration.plus_z(p)
    chain.jump_to_image(p, 0)
    transition.homogeneous(p, 0, noi - 1)
    simulation.start(p, simulation.METHOD_GNEB, simulation.SOLVER_VP, n_iterations=200)
    gneb.set_image_type_automatically(p)
    t0 = time.perf_counter()
    info = simulation.start(p, simulation.METHOD_GNEB, simulation.SOLVER_VP, n_iterations=n_iter)
    t = time.perf_counter() - t0
    E = np.array(chain.get_energy(p))
    nos = 128 * 128
    state.delete(p)
    value = nos * noi * info.total_iterations / t
    ach = GNEB_VP_BYTES_PER_SPIN_ITERATION * value / 1e9
    return {"workload": "hex128_gneb%d_vp_climbing" % noi, "images": noi, "spins_per_image": nos,
            "iterations": info.total_iterations, "seconds": t, "iterations_per_s": info.total_iterations / t,
            "value": value, "unit": "spin-iterations/s",
            "roofline": {"bound": "hbm", "alg_bytes_per_spin_iteration": GNEB_VP_BYTES_PER_SPIN_ITERATION,
                         "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                         "timed": "the whole Simulation_GNEB_Start call (set-up, copies, run)"},
            "energy_first_last_max_meV": [float(E[0]), float(E[-1]), float(E.max())],
            "max_torque": float(info.max_torque)}


def _cube_state(cfg_over, site_K=True):
    """State on the default 256x256x128 lattice (overrides applied), per-site K_i and
    random spins from the workload's seeds."""
    from spirit_amd import configs, hip, state, system
    cfg = dict(configs.CUBE_256, **cfg_over)
    nos = 256 * 256 * 128
    path = os.path.join(tempfile.mkdtemp(), "cube.cfg")
    configs.write_config(cfg, path)
    p = state.setup(path, True)
    if not p:
        raise RuntimeError("State_Setup failed")
    if site_K:
        hip.set_site_anisotropy(p, configs.site_anisotropy(nos, 20006), (0.0, 0.0, 1.0))
    system.get_spin_directions(p)[:] = configs.random_spins(nos, 20006)
    return p, nos


def run_minimiser(n_iter, warmup=20):
    """LBFGS_OSO direct minimisation of the default workload (DDI included): n_iter
    iterations (one force evaluation plus the L-BFGS update each) through n_shot."""
    from spirit_amd import hip, simulation, state
    p, nos = _cube_state({})
    hip.set_host_sync(p, 0)
    simulation.start(p, simulation.METHOD_LLG, simulation.SOLVER_LBFGS_OSO, single_shot=True)
    simulation.n_shot(p, warmup)
    t0 = time.perf_counter()
    simulation.n_shot(p, n_iter)
    t = time.perf_counter() - t0
    torque = float(simulation.get_max_torque_norm(p))
    simulation.stop(p)
    state.delete(p)
    return {"workload": "cube256_lbfgs_oso_ddi", "spins": nos, "iterations": n_iter, "seconds": t,
            "ms_per_iteration": 1e3 * t / n_iter, "value": nos * n_iter / t, "unit": "spin-iterations/s",
            "max_torque_after": torque}


def run_mc(n_short, n_long):
    """Metropolis Monte Carlo (adaptive cone, T = 10 K) on the 256x256x128 lattice without
    DDI (the single-spin energy has none): the rate is taken from two runs of n_short and
    n_long sweeps, so set-up and copies cancel."""
    from spirit_amd import simulation, state
    from spirit_amd.parameters import mc
    p, nos = _cube_state({"ddi_method": "none"})
    mc.set_temperature(p, 10.0)
    mc.set_metropolis_cone(p, True, 30.0, True, 0.5)
    times = []
    for n in (n_short, n_long):
        t0 = time.perf_counter()
        info = simulation.start(p, simulation.METHOD_MC, n_iterations=n)
        times.append(time.perf_counter() - t0)
        if info.total_iterations != n:
            raise RuntimeError("MC ran %d of %d sweeps" % (info.total_iterations, n))
    state.delete(p)
    dt = times[1] - times[0]
    return {"workload": "cube256_metropolis_T10", "spins": nos, "sweeps": n_long - n_short, "seconds": dt,
            "ms_per_sweep": 1e3 * dt / (n_long - n_short), "value": nos * (n_long - n_short) / dt,
            "unit": "spin-sweeps/s", "call_seconds": times}


PASS_C_KERNELS = {0: "k_pass_c_conv", 1: "k_p2_c", 2: "k_p2_c1", 3: "k_p4_c1"}


def pass_c_bytes(dims, info):
    """HBM bytes one convolution-pass launch must move: the c < Nc planes of the spin
    spectrum read and written (96 H_PN) plus one read of the kernel spectrum as stored
    (8 B per real entry, 16 B per complex one; 6 tensor components)."""
    import re
    Na, Nb, Nc = dims
    Pb = 2 * Nb if Nb > 1 else 1
    Ha = (2 * Na if Na > 1 else 1) // 2 + 1
    m = re.search(r"stored (\d+)x(\d+)x(\d+)", info)
    kb, ha, kc = (int(x) for x in m.groups())
    entry = 8 if "spectrum real" in info else 16
    kind = int(re.search(r"pass C kernel (\d+)", info).group(1))
    return 96 * Ha * Pb * Nc + 6 * entry * kb * ha * kc, PASS_C_KERNELS.get(kind, str(kind))


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE, scaled by a copy of known size); None if not recorded."""
    import re
    try:
        text = open(os.path.join(ROOT, "profiles", "hbm_traffic_cube256.md")).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch: %s = (\d+)" % re.escape(kernel), text)
    return int(m.group(1)) if m else None


def cpu_baseline(wl, threads, steps, fft_parallel=True):
    """The CPU oracle (C + OpenMP restatement of the reference algorithm) on the host cores,
    timed on `steps` steps of the workload `wl` after its set-up.  fft_parallel=False runs
    each 3-D transform on one thread, as the reference's kissFFT build does (the rest of the
    step still uses every thread, as the reference's OpenMP loops do)."""
    code = r"""
import sys, time, json
sys.path[:0] = [%r, %r]
import oracle as O
from spirit_amd import configs
O.set_fft_parallel(%r)
cfg = dict(getattr(configs, %r)); n = %d
K = configs.site_anisotropy(n, 20006) if %r else None
S = O.System(cfg, site_K=K)
s = configs.random_spins(n, 20006)
t0 = time.perf_counter()
s, e, t, _ = S.llg(s, %d, %d, dt=cfg["llg_dt"], damping=cfg["llg_damping"], temperature=cfg["llg_temperature"], seed=cfg["llg_seed"])
print(json.dumps({"seconds": time.perf_counter() - t0}))
""" % (ROOT, os.path.join(ROOT, "oracle"), bool(fft_parallel), wl["cfg"], wl["dims"][0] * wl["dims"][1] * wl["dims"][2],
       wl["site_K"], wl["solver"], steps)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=1200)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])["seconds"]


def host_threads():
    """Threads for the CPU baseline: the CPUs this process may run on, at most 16 (the host's
    share for one GPU), and the machine's count as reported."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return min(16, usable), usable, os.cpu_count() or 1


def cpu_baselines(main_wl, main_steps, threads):
    """The CPU baselines, each timed on the host cores with `threads` OpenMP threads:
    the headline workload (main_steps steps, threaded FFT: a reference-faithful serial-FFT step
    of the 512x512x256 grid takes over a minute); the 64x64 Depondt monolayer (2000 steps) and
    the 128x128x32 Heun+DDI film (1 step) with each transform on one thread, as the reference's
    kissFFT build runs them (the reference-faithful baseline), and the film once more with the
    threaded FFT."""
    out = {}
    sec = cpu_baseline(main_wl, threads, main_steps, True)
    n = main_wl["dims"][0] * main_wl["dims"][1] * main_wl["dims"][2]
    out["main"] = (n * main_steps / sec, sec)
    for key, wl, steps, par in (("square64_depondt_2000", WORKLOADS["square64_depondt"], 2000, False),
                                ("film128_heun_ddi_serial_fft", WORKLOADS["film128_heun_ddi"], 1, False),
                                ("film128_heun_ddi_threaded_fft", WORKLOADS["film128_heun_ddi"], 1, True)):
        n = wl["dims"][0] * wl["dims"][1] * wl["dims"][2]
        sec = cpu_baseline(wl, threads, steps, par)
        out[key] = {"value": n * steps / sec, "unit": "spin-steps/s", "steps": steps, "seconds": sec,
                    "fft": "threaded" if par else "serial per transform (reference-faithful)"}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cube256_depondt_ddi", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs", default="plain,batched,adjoint,gneb,minimiser,mc",
                    help="comma-separated: plain (always run), batched (uneven ensemble), adjoint, gneb, "
                         "minimiser (LBFGS_OSO), mc (Metropolis)")
    ap.add_argument("--minimiser-iterations", type=int, default=200)
    ap.add_argument("--mc-sweeps", type=int, default=1000)
    ap.add_argument("--gneb-iterations", type=int, default=2000)
    ap.add_argument("--gneb-images", type=int, default=32)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-ensemble-sequential", action="store_true",
                    help="skip the one-replica-at-a-time comparison run (profiling the batched run alone)")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--adjoint-checkpoint", type=int, default=50)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device three times (a known-size copy kernel "
                         "for scaling hardware byte counters)")
    ap.add_argument("--copy-rates", action="store_true",
                    help="print the device's copy / read / write bandwidth (one JSON line) and exit")
    args = ap.parse_args()
    if args.copy_rates:
        print(json.dumps(copy_rates()), flush=True)
        return
    legs = set(args.legs.split(","))
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (use --gpus 1)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    ddi = args.workload != "square64_depondt"
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
    g = run_gpu(args, wl)
    value = g["nos"] * args.steps / g["seconds"]
    b_alg, conv_bytes = alg_bytes_per_spin_step(wl["dims"], wl["solver"], ddi)
    ms_per_step = 1e3 * g["seconds"] / args.steps
    out = {
        "metric": "LLG spin-steps/s (Depondt+DDI) plain vs batched vs adjoint; % HBM peak",
        "value": value,
        "unit": "spin-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random spins and per-site anisotropy)",
        "config": {"workload": args.workload, "lattice": list(wl["dims"]), "solver": {2: "Depondt", 3: "Heun"}.get(
            wl["solver"], str(wl["solver"])), "ddi": "fft" if ddi else "none", "spins": g["nos"],
            "alg_bytes_per_spin_step": b_alg, "hbm_frac_of_step": b_alg * value / (HBM_PEAK_GBS * 1e9)},
    }
    if g["conv_n"] > 0:
        t = g["conv_ms"] / g["conv_n"] * 1e-3
        c_bytes, c_kernel = pass_c_bytes(wl["dims"], g["ddi_info"])
        d_bytes = design_bytes_per_spin_step(wl["dims"], wl["solver"], c_bytes)
        out["config"]["design_bytes_per_spin_step"] = d_bytes
        out["config"]["hbm_frac_of_step_design_bytes"] = d_bytes * value / (HBM_PEAK_GBS * 1e9)
        ach = c_bytes / t / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS,
                           "traffic": measured_traffic(c_kernel) if args.workload == "cube256_depondt_ddi" else None,
                           "kernel": c_kernel,
                           "alg_bytes_per_launch": c_bytes, "dense_spectrum_bytes_per_launch": conv_bytes,
                           "avg_launch_ms": t * 1e3}
        out["config"]["ddi_plan"] = g["ddi_info"]
    else:
        out["roofline"] = None
    out["kernel_ms"] = {"ddi_conv_avg": g["conv_ms"] / max(1, g["conv_n"]),
                        "llg_stage_avg": g["stage_ms"] / max(1, g["stage_n"])}
    out["max_norm_error"] = g["norm_err"]
    if "batched" in legs:
        out["batched"] = run_batched(args.ensemble_steps, args.replicas, not args.no_ensemble_sequential)
    if "adjoint" in legs:
        out["adjoint"] = run_adjoint(args.adjoint_steps, args.adjoint_checkpoint,
                                     ms_per_step if args.workload == "cube256_depondt_ddi" else None)
    if "gneb" in legs:
        out["gneb"] = run_gneb(args.gneb_iterations, args.gneb_images)
    if "minimiser" in legs:
        out["minimiser"] = run_minimiser(args.minimiser_iterations)
    if "mc" in legs:
        out["mc"] = run_mc(max(1, args.mc_sweeps // 5), args.mc_sweeps + max(1, args.mc_sweeps // 5))
    if not args.no_cpu_baseline:
        threads, usable, nproc = host_threads()
        try:
            cb = cpu_baselines(wl, args.cpu_steps, threads)
            value, _ = cb.pop("main")
            out["cpu_baseline"] = {
                "value": value, "unit": "spin-steps/s", "cores": threads, "kind": "port",
                "sample": "%d step(s) of the same workload after set-up, FFT lines threaded; %d OpenMP threads "
                          "(OMP_PROC_BIND=close) of %d usable CPUs (%d reported), CPU %s" % (
                              args.cpu_steps, threads, usable, nproc, cpu_model()),
                "reference_faithful": cb,
                "reference_faithful_note": "the oracle with each FFT on one thread, as the reference's kissFFT "
                                           "build runs its transforms; the baseline for the 64x64 and 128x128x32 "
                                           "configurations"}
        except Exception as e:  # reported, never fatal for the GPU number
            out["cpu_baseline"] = {"value": None, "unit": "spin-steps/s", "cores": threads, "kind": "port",
                                   "sample": "failed: %s" % str(e)[-300:]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
