"""Throughput of the WCSPH step on one MI355X: Taylor-Green vortex on a 2048^2
lattice (4,194,304 fluid particles + 32,832 periodic ghosts), predictor-corrector
(time_iteration.f90, leapforg_or_eular=2). Both forms run on the same workload:
the plain FP64 form and the FP32-mixed form (FP32 pair sums with Kahan
accumulation, FP64 positions, integration and EOS); `value` is the mixed form's
particle-steps/s, with the FP64 figure, the speedup and the KE error between the
two forms in `config`.

One step = one pass of time_iteration.f90's loop body: periodic ghosts, cell-key
sort + reorder, stage updates with one force evaluation (two on step 1), MLS
density reinitialisation every 20 steps, particle shifting and the refresh.
The state is resident in HBM when the timed region starts; timing uses HIP
events on the context's stream. Prints ONE JSON line.

Also reported: the HBM roofline of the dominant neighbour pass (the frozen
algorithmic bytes per particle of that pass times the particles of one launch,
over its duration from HIP events recorded around the pass inside real steps on
the context's stream), the same model for the whole step, the kinetic-energy
error between the two forms at t = 0.5, and the CPU baseline (the C oracle on
the host: one core on the 2048^2 workload itself, and one process per core of
the box's share on 512^2 samples).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sph-2d_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
FP64_VECTOR_TFLOPS = 78.6


def hip_events():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return hip


# Frozen algorithmic HBM bytes per fluid particle (compulsory traffic of the fused
# three-pass design: each per-particle array read once and written once per pass,
# neighbour data served from LDS/L2, pairs never stored; FP64 state x, y, u, v, rho,
# m at 8 B plus type 4 B = 52 B, the mixed form's u, v, delta rho, m at 4 B = 36 B).
PASS_BYTES = {
    # P1 neighbour sums (S_xx, S_xy, S_yy, grad rho, fra_v, sum dW): read state, write 8 values
    "pass_a": {64: 52 + 64, 32: 36 + 32},
    # P3 forces: read state + S + grad rho + flag, write drho/dt and a
    "force": {64: 93 + 24, 32: 57 + 12},
}
# one predictor-corrector step with one evaluation (P1 + P2 flags + P3), the stage
# update, the key/sort/permute and the shift: 295 + 144 + 168 + 76 (FP64) and
# 179 + 96 + 136 + 56 (mixed) bytes per particle-step
STEP_BYTES = {64: 295 + 144 + 168 + 76, 32: 179 + 96 + 136 + 56}


# HBM bytes per launch of the mixed passes at 2048^2, per subdomain particle, measured by
# rocprofv3 --pmc FETCH_SIZE and then WRITE_SIZE on this build, each counter scaled by
# what it reads for a 256 MiB device-to-device copy (summary and commands in the source).
_TRAFFIC_SRC = "profiles/tgv2048_force_pass_hbm_traffic.md"
PASS_TRAFFIC = {
    "force": {"bytes_per_particle": 151.8, "n_side": 2048, "source": _TRAFFIC_SRC},
    "pass_a": {"bytes_per_particle": 163.9, "n_side": 2048, "source": _TRAFFIC_SRC},
}
# what each mixed pass reads and writes for its own particle, per launch:
# force pass: records 48 + (p, c) 8, type, cell, near 12, neighbour count 4 and list ~40;
#   writes drho/dt 8, acceleration 16, shift 16
# pass A: r0 16, FP64 position 16, FP32 position 8, type and cell 8; writes the list ~36,
#   its count 4, fra_v 8, sum dW 16, tau 16 (r1), grad rho 16 (r2), flags 8
PASS_READ_WRITE = {"force": 48 + 8 + 12 + 44 + 40, "pass_a": 48 + 104}
# FLOPs per subdomain particle of one launch (SQ_INSTS_VALU_FLOPS_* x 64 lanes)
_VALU_SRC = "profiles/tgv2048_force_pass_valu.md"
PASS_VALU = {
    "force": {"fp32_flop_per_particle": 5293.0, "fp64_flop_per_particle": 2.0, "source": _VALU_SRC},
    "pass_a": {"fp32_flop_per_particle": 3094.0, "fp64_flop_per_particle": 225.0, "source": _VALU_SRC},
}
FP32_VECTOR_TFLOPS = 157.3
FP32_VECTOR_TFLOPS_UNPACKED = 78.6


def cpu_baseline(n_side: int = 1024, steps: int = 1) -> dict:
    """The C oracle (oracle/liboracle.so, the reference's algorithm restated,
    scatter over a half pair list as in find_particle_pair.f90), one core, on a
    bounded sample of the workload itself: n_side^2 TGV, predictor-corrector,
    `steps` steps after step 1 (which has the extra evaluation)."""
    from oracle.pyoracle import Oracle
    o = Oracle(nx=n_side, ny=n_side, integrator=2, dt=1e-4 * 100.0 / n_side)
    o.init_tgv()
    o.time_step(1)
    t0 = time.perf_counter()
    for it in range(2, 2 + steps):
        o.time_step(it)
    dt = time.perf_counter() - t0
    o.close()
    n = n_side * n_side
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"TGV {n_side}^2 ({n} particles), predictor-corrector FP64, steps 2..{1 + steps} "
                      f"timed ({dt:.1f} s); oracle/liboracle.so single-threaded"}


_CPU_WORKER = """
import sys, time
sys.path.insert(0, {root!r})
from oracle.pyoracle import Oracle
n, steps = {n}, {steps}
o = Oracle(nx=n, ny=n, integrator=2, dt=1e-4 * 100.0 / n)
o.init_tgv()
o.time_step(1)
t0 = time.perf_counter()
for it in range(2, 2 + steps):
    o.time_step(it)
print(n * n * steps / (time.perf_counter() - t0))
"""

# host cores a one-GPU job may use on these machines (its share of the node, whatever
# lscpu reports for the whole node)
BOX_CORE_SHARE = 16


def physical_cores() -> int:
    """Physical cores of the host from lscpu (sockets x cores per socket)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        return int(kv["Socket(s)"].strip()) * int(kv["Core(s) per socket"].strip())
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline_all_cores(n_side: int, steps: int) -> dict:
    """One oracle process per core (the oracle, like one reference rank, runs one
    simulation on one core), each its own instance of an n_side^2 sample, all at
    once; the value is the sum of their rates. The core count is lscpu's physical
    cores, capped at the host share of a one-GPU job (16 cores; a 1024^2 instance
    holds ~4.4 GB, so 16 of them fit the job's host memory where 2048^2 ones would
    not). Children are started as new programs, not forks of this GPU process."""
    import subprocess
    phys = physical_cores()
    cores = max(1, min(phys, BOX_CORE_SHARE))
    code = _CPU_WORKER.format(root=ROOT, n=n_side, steps=steps)
    t0 = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, text=True) for _ in range(cores)]
    rates = []
    for p in procs:
        out, _ = p.communicate(timeout=900)
        if p.returncode == 0:
            rates.append(float(out.strip().splitlines()[-1]))
    return {"value": sum(rates), "unit": "particle-steps/s", "cores": len(rates),
            "physical_cores_lscpu": phys, "core_share": BOX_CORE_SHARE,
            "kind": "port", "per_core_mean": sum(rates) / max(1, len(rates)),
            "wall_s": time.perf_counter() - t0,
            "sample": f"{len(rates)} concurrent single-core oracle processes (independent instances), each "
                      f"TGV {n_side}^2 predictor-corrector FP64, steps 2..{1 + steps} timed"}


def ke_gate_leg(sph2d, n_side: int, t_end: float = 0.5) -> dict:
    """The accuracy bar of the mixed form: both forms run the n_side^2 TGV
    (predictor-corrector, dt = 1e-4 * 100 / n_side) to t_end from the same initial
    condition; relative KE/KE0 difference and each form's error against the analytic
    decay exp(-16 pi^2 nu t)."""
    import math
    dt = 1e-4 * 100.0 / n_side
    steps = int(round(t_end / dt))
    res = {}
    for prec in (sph2d.FP32_MIXED, sph2d.FP64):
        sim = sph2d.Simulation(sph2d.tgv_params(n_side, integrator=sph2d.INTEGRATOR_PC, precision=prec))
        sim.init_tgv()
        t0 = time.perf_counter()
        chunk = 5000
        for first in range(1, steps + 1, chunk):
            sim.run(first, min(chunk, steps + 1 - first))
            sim.diagnostics()                      # synchronises: a progress line per chunk
            print(f"# ke gate form {prec}: step {min(first + chunk - 1, steps)} of {steps}", file=sys.stderr,
                  flush=True)
        s = sim.sample()
        res[prec] = (s, time.perf_counter() - t0)
        sim.close()
    km, kf = res[sph2d.FP32_MIXED][0][0], res[sph2d.FP64][0][0]
    nu = sph2d.tgv_params(n_side).nu
    ka = math.exp(-16 * math.pi ** 2 * nu * t_end)
    return {"n_side": n_side, "t": steps * dt, "steps": steps, "ke_over_ke0_mixed": km, "ke_over_ke0_fp64": kf,
            "ke_rel_err_mixed_vs_fp64": abs(km - kf) / kf, "bar": 1e-4,
            "ke_analytic": ka, "fp64_vs_analytic": (kf - ka) / ka, "mixed_vs_analytic": (km - ka) / ka,
            "l2_u_mixed": res[sph2d.FP32_MIXED][0][4], "l2_u_fp64": res[sph2d.FP64][0][4],
            "wall_s": {"mixed": res[sph2d.FP32_MIXED][1], "fp64": res[sph2d.FP64][1]}}


def batched_leg(sph2d, instances: int, steps: int, warmup: int, trial: int = 20) -> dict:
    """Many independent periodic TGV instances of uneven size and Reynolds number
    (n_side = round(64 * 8^U), Re = 50 * 20^U', U, U' uniform, seed 3), FP32-mixed
    predictor-corrector: one sph2d_batch run (each replayable step one hipGraph
    launch) against the same instances run one after another, each alone."""
    import numpy as np
    sides, res = sph2d.uneven_tgv_instances(instances, seed=3)
    total = sum(n * n for n in sides)

    def make():
        sims = []
        for n, re in zip(sides, res):
            sim = sph2d.Simulation(sph2d.tgv_params(n, integrator=sph2d.INTEGRATOR_PC, precision=sph2d.FP32_MIXED,
                                                    nu=1.0 / re))
            sim.init_tgv()
            sims.append(sim)
        return sims

    hip = hip_events()
    sims = make()
    batch = sph2d.Batch(sims)
    batch.run(1, warmup)
    # the autotuner times each launch configuration on `trial` real steps (20: one MLS
    # step each); trial 0 keeps configuration 0
    if trial > 0:
        tune_ms, chosen = batch.autotune(1 + warmup, trial)
    else:
        tune_ms, chosen = np.zeros(len(sph2d.Batch.CONFIGS)), 0
    tried = int(np.count_nonzero(tune_ms))
    first = 1 + warmup + tried * (trial + 1)
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    batch.run(first, steps)
    hip.hipDeviceSynchronize()
    t_batch = time.perf_counter() - t0
    nodes, replays = batch.info()
    ke = [s.sample()[0] for s in sims]
    batch.close()
    for s in sims:
        s.close()
    sims = make()
    for s in sims:
        s.run(1, first - 1)
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for s in sims:
        s.run(first, steps)
        hip.hipDeviceSynchronize()
    t_seq = time.perf_counter() - t0
    same = all(abs(s.sample()[0] - k) == 0.0 for s, k in zip(sims, ke))
    for s in sims:
        s.close()
    print(f"# batched: {instances} instances, {total} particles, {steps} steps: graph {t_batch:.3f} s, "
          f"one after another {t_seq:.3f} s", file=sys.stderr, flush=True)
    return {"instances": instances, "particles": total, "n_side_min": min(sides), "n_side_max": max(sides),
            "re_min": min(res), "re_max": max(res), "steps": steps, "form": "FP32-mixed, predictor-corrector",
            "particle_steps_per_s": total * steps / t_batch, "ms_per_step": 1000.0 * t_batch / steps,
            "sequential_particle_steps_per_s": total * steps / t_seq,
            "speedup_vs_sequential": t_seq / t_batch, "graph_nodes": nodes, "graph_replays": replays,
            "autotune": {"ms_per_step": {name: float(v) for name, v in zip(sph2d.Batch.CONFIGS, tune_ms)},
                         "chosen": sph2d.Batch.CONFIGS[chosen], "trial_steps": trial,
                         "default_over_chosen": float(tune_ms[0] / tune_ms[chosen]) if tried else None,
                         "slowest_over_chosen": float(tune_ms.max() / tune_ms[chosen]) if tried else None},
            "timed_steps_from": first,
            "ke_identical_to_sequential": same,
            "timing": "host wall clock with device synchronisation, launches included"}


def evaluation_vjp_leg(sph2d, n_side: int, reps: int) -> dict:
    """Reverse of one FP64 force evaluation (sph2d_single_step_compute_vjp) beside the
    forward evaluation it reverses (sph2d_single_step_compute), TGV lattice of n_side^2:
    wall-clock per call between device synchronisations. The reverse call includes
    its host-side weights upload and the gradient download (its device buffers are
    allocated by the untimed first call); the forward call is device work only."""
    import numpy as np
    sim = sph2d.Simulation(sph2d.tgv_params(n_side, integrator=2))
    sim.init_tgv()
    sim.distribute_period_boundary()
    sim.devide_block(1)
    hip = sph2d._hiprt()
    sim.single_step_compute(1)
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.single_step_compute(1)
    hip.hipDeviceSynchronize()
    fwd = (time.perf_counter() - t0) / reps
    N = n_side * n_side
    rng = np.random.default_rng(0)
    lam = [rng.normal(0.0, 1.0, N) for _ in range(3)]
    sim.single_step_compute_vjp(*lam)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.single_step_compute_vjp(*lam)
    rev = (time.perf_counter() - t0) / reps
    # the same reverse on device arrays (sph2d_single_step_compute_vjp_device): the
    # kernels only, no host packing or transfers
    d_ld, d_la = sph2d.DeviceArray(lam[0]), sph2d.DeviceArray(np.concatenate([lam[1], lam[2]]))
    gp, gv, gr = (sph2d.DeviceArray(np.zeros(k * N)) for k in (2, 2, 1))
    sim.single_step_compute_vjp_device(d_ld, d_la, gp, gv, gr)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.single_step_compute_vjp_device(d_ld, d_la, gp, gv, gr)
    rev_d = (time.perf_counter() - t0) / reps
    for a in (d_ld, d_la, gp, gv, gr):
        a.free()
    sim.close()
    return {"n_side": n_side, "particles": N, "forward_evaluation_ms": 1e3 * fwd,
            "reverse_evaluation_ms": 1e3 * rev, "reverse_over_forward": rev / fwd,
            "reverse_evaluation_device_ms": 1e3 * rev_d, "reverse_device_over_forward": rev_d / fwd, "reps": reps}


# the wave tank of the adjoint workload: 8 m long, water 0.4 m deep, dx = 0.005 m
TANK = dict(length=8.0, height=0.6, depth=0.4, dx=0.005, g=9.81, damping_length=2.0)


def wave_tank(sph2d, dx: float = TANK["dx"]):
    """(params, fluid state) of the piston wave tank: still water at hydrostatic
    pressure (rho from the inverse Tait law with the REAL(4) exponent), c0 = 10
    sqrt(g depth), rho0 = 1000, walls of 3 layers, 3 wavemaker columns, a linear
    damping zone over the last 2 m, MLS reinitialisation off (the reverse holds no
    MLS), dt = 0.083 h / c0 (the reference's c0 dt / h)."""
    import numpy as np
    L, Hb, d, g = TANK["length"], TANK["height"], TANK["depth"], TANK["g"]
    nx, ny = int(round(L / dx)), int(round(Hb / dx))
    c0 = 10.0 * np.sqrt(g * d)
    h = 1.2 * dx
    prm = sph2d.default_params(nx=nx, ny=ny, box_x=L, box_y=Hb, dx=dx, h=h, max_shift=0.25 * dx,
                               rho0=1000.0, c0sq=c0 * c0, nu=1e-6, U=float(np.sqrt(g * d)),
                               dt=0.083 * h / c0, integrator=2, walls=3, periodic=0, gravity=g,
                               wavemaker_layers=3, damping=sph2d.DAMPING_LINEAR,
                               damping_length=TANK["damping_length"], damping_amplitude=0.5, water_depth=d,
                               mls_on=0)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(int(round(d / dx))), indexing="ij")
    x = ((ii + 0.5) * dx).reshape(-1)
    y = ((jj + 0.5) * (Hb / ny)).reshape(-1)
    p = 1000.0 * g * (d - y)
    rho = 1000.0 * (1 + 7 * p / (c0 * c0 * 1000.0)) ** float(np.float32(0.142857142))
    M = len(x)
    return prm, (x, y, np.zeros(M), np.zeros(M), rho, rho * dx * (Hb / ny))


def adjoint_leg(sph2d, steps: int, k: int = 20) -> dict:
    """The adjoint workload on the full tank (~128k fluid particles): sph2d_forward over
    `steps` steps (8-harmonic stroke, f0 = 0.5 Hz, a_k ~ N(0, 0.01 m), seed 4; MLS
    pressure probes at x = 2, 3, 4 m, 0.1 m below the still surface, every 10 steps;
    a checkpoint every k steps), then sph2d_backward of a misfit against a second set
    of coefficients: wall clock of each between device synchronisations."""
    import numpy as np
    prm, st = wave_tank(sph2d)
    rng = np.random.default_rng(4)
    a, a_t = rng.normal(0.0, 0.01, 8), rng.normal(0.0, 0.01, 8)
    cfg = sph2d.adjoint_cfg(n_harmonics=8, f0=0.5, standing_steps=1, sample_every=10,
                            probes=[(2.0, 0.3), (3.0, 0.3), (4.0, 0.3)], checkpoint_every=k)
    hip = sph2d._hiprt()

    def fresh():
        sim = sph2d.Simulation(prm, capacity=len(st[0]))
        sim.upload_state(*st)
        sim.wavemaker_update(0.0, 0.0)
        sim.time_step(1)
        return sim

    sim = fresh()
    target = sim.forward(cfg, a_t, 2, steps)
    sim.close()
    sim = fresh()
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    p = sim.forward(cfg, a, 2, steps)
    tf = time.perf_counter() - t0
    w = 2.0 * (p - target) / p.size
    t0 = time.perf_counter()
    g = sim.backward(w)
    tb = time.perf_counter() - t0
    n_fluid = sim.n_fluid
    sim.close()
    return {"workload": "piston wave tank 8 m x 0.4 m deep, dx 0.005 m, FP64, predictor-corrector",
            "fluid_particles": n_fluid, "steps": steps, "dt": prm.dt, "checkpoint_every": k,
            "steps_for_2s": int(round(2.0 / prm.dt)),
            "forward_ms_per_step": 1e3 * tf / steps, "backward_ms_per_step": 1e3 * tb / steps,
            "reverse_over_forward": tb / tf, "loss": float(np.mean((p - target) ** 2)),
            "grad_norm": float(np.linalg.norm(g))}


def moving_square_leg(sph2d, steps: int, res, every: int = 100) -> dict:
    """The moving square (the reference's Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box
    program): 10 D x 5 D tank, dx = D/100 (490,000 fluid particles, 3 ghost layers), the
    square ramped to U = 1 over t = 0.5, MUSCL-HLLC + delta-SPH + SPS + shifting, the FP64
    and FP32-mixed forms for each Re. Reports each form's particle-steps/s (device time of
    the step loop, body-force downloads excluded) and the force on the water every `every`
    steps (in units of 0.5 rho0 U^2 D), and the mixed form's drag and lift against FP64."""
    import numpy as np
    hip = hip_events()
    out = {"workload": f"moving square, 10D x 5D tank, dx = D/100, {steps} steps per form, MUSCL-HLLC + "
                       f"delta-SPH + SPS + shifting, MLS order 3 every 40 steps", "steps": steps, "cases": {}}
    for re_ in res:
        case = {}
        for prec, name in ((sph2d.FP64, "fp64"), (sph2d.FP32_MIXED, "mixed")):
            p = sph2d.moving_square_params(n_per_d=100, re=float(re_), precision=prec, sample_every=every)
            sim = sph2d.Simulation(p, capacity=p.nx * p.ny)
            sim.init_moving_square()
            stream = ctypes.c_void_p(sim.stream())
            e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
            hip.hipEventCreate(ctypes.byref(e0))
            hip.hipEventCreate(ctypes.byref(e1))
            hist, dev_ms = [], 0.0
            it = 1
            while it <= steps:
                k = min(every - (it - 1) % every, steps - it + 1)
                hip.hipEventRecord(e0, stream)
                sim.run(it, k)
                hip.hipEventRecord(e1, stream)
                hip.hipEventSynchronize(e1)
                ms = ctypes.c_float()
                hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
                dev_ms += ms.value
                it += k
                if (it - 1) % every == 0:
                    bf, st = sim.body_force()
                    hist.append([st * p.dt] + [float(v) / 0.5 for v in bf])
            hip.hipEventDestroy(e0)
            hip.hipEventDestroy(e1)
            d = sim.diagnostics()
            g = sim.download_state()
            case[name] = {"particle_steps_per_s": sim.n_fluid * steps / (dev_ms * 1e-3),
                          "ms_per_step": dev_ms / steps, "n_local": d.n_local,
                          "history": hist, "finite": bool(np.all(np.isfinite(g["u"])) and np.all(np.isfinite(g["v"]))),
                          "error_flags": d.error_flags, "rebin_needed": d.rebin_needed}
            sim.close()
            print(f"# moving square Re {re_} {name}: {dev_ms / steps:.3f} ms/step", file=sys.stderr, flush=True)
        h64 = np.array(case["fp64"]["history"])
        h32 = np.array(case["mixed"]["history"])
        drag64, drag32 = h64[:, 5], h32[:, 5]
        lift64, lift32 = h64[:, 6], h32[:, 6]
        scale = np.max(np.abs(drag64))
        case["t_end"] = float(h64[-1, 0])
        case["total_drag_coefficient_end"] = {"fp64": float(drag64[-1]), "mixed": float(drag32[-1])}
        case["total_lift_coefficient_end"] = {"fp64": float(lift64[-1]), "mixed": float(lift32[-1])}
        case["mixed_vs_fp64_drag_max_rel"] = float(np.max(np.abs(drag32 - drag64)) / scale)
        case["mixed_vs_fp64_lift_max_rel_to_drag"] = float(np.max(np.abs(lift32 - lift64)) / scale)
        case["mixed_over_fp64_speedup"] = case["mixed"]["particle_steps_per_s"] / case["fp64"]["particle_steps_per_s"]
        out["cases"][str(re_)] = case
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n-side", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-all-side", type=int, default=1024,
                    help="CPU baseline (the reported value): one oracle instance of this lattice per core")
    ap.add_argument("--cpu-all-steps", type=int, default=1)
    ap.add_argument("--cpu-side", type=int, default=1024, help="one-core CPU sample on this lattice")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-small-side", type=int, default=100,
                    help="all-core CPU sample on the reference's own 100^2 case")
    ap.add_argument("--cpu-small-steps", type=int, default=500)
    ap.add_argument("--ke-gate-side", type=int, default=1024,
                    help="KE error of the mixed form vs FP64 at t = 0.5 on this lattice (0 skips; 2048 is the "
                         "4.2M-particle gate, 102,400 steps per form)")
    ap.add_argument("--forms", choices=("both", "fp64", "mixed", "batched", "ms", "adjoint"), default="both",
                    help="both: the full line; fp64/mixed: that form's timing only; batched: the batched "
                         "leg only; ms: the moving-square leg only; adjoint: the wave-tank adjoint leg only")
    ap.add_argument("--timing-steps", type=int, default=20,
                    help="steps after the timed run that record the per-pass HIP events")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB: a plain copy kernel of "
                         "known size that scales rocprofv3's FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--batch-instances", type=int, default=64,
                    help="batched leg: independent TGV instances advanced by sph2d_batch (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-trial-steps", type=int, default=20,
                    help="autotuner steps per launch configuration (0 keeps configuration 0)")
    ap.add_argument("--adjoint-side", type=int, default=1024,
                    help="reverse of one evaluation on an n^2 TGV lattice (0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="wave-tank adjoint leg: steps of sph2d_forward/sph2d_backward (0 skips it)")
    ap.add_argument("--ms-steps", type=int, default=5000,
                    help="moving-square leg: steps per form and Reynolds number (0 skips it)")
    ap.add_argument("--ms-re", type=str, default="50,100,150", help="moving-square leg: Reynolds numbers")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU form (spatial slabs "
                                   "with a halo exchange per evaluation) is designed but not built"}))
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import sph2d
    n = args.n_side
    hip = hip_events()

    def run_form(precision: int) -> dict:
        prm = sph2d.tgv_params(n, integrator=sph2d.INTEGRATOR_PC, precision=precision)
        sim = sph2d.Simulation(prm)
        sim.init_tgv()
        if args.warmup > 0:
            sim.run(1, args.warmup)
        sim.diagnostics()                          # synchronises
        stream = ctypes.c_void_p(sim.stream())
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        hip.hipEventRecord(e0, stream)
        sim.run(1 + args.warmup, args.steps)
        hip.hipEventRecord(e1, stream)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        total_s = ms.value / 1000.0
        smp = sim.sample()
        d = sim.diagnostics()
        if d.error_flags or d.rebin_needed:
            print(f"# warning: error_flags={d.error_flags} rebin={d.rebin_needed}", file=sys.stderr)
        # pass durations inside real steps (HIP events around the passes), in a short
        # run after the timed one so the events do not perturb it
        sim.set_pass_timing(True)
        sim.run(1 + args.warmup + args.steps, max(1, args.timing_steps))
        force_ms, passa_ms, eval_ms, n_eval = sim.pass_timing()
        sim.set_pass_timing(False)
        cnt = ctypes.c_int64()
        sph2d._check(sph2d.lib().sph2d_download_pairs(sim._c, 0, None, None, ctypes.byref(cnt)))
        out = {"total_s": total_s, "sample": smp, "n_local": d.n_local, "n_ghost": d.n_ghost,
               "force_ms": force_ms, "passa_ms": passa_ms, "eval_ms": eval_ms, "n_eval_timed": n_eval,
               "pairs": cnt.value,
               "dt": prm.dt}
        sim.close()
        print(f"# form {precision}: {args.steps} steps in {total_s:.3f} s, KE/KE0={smp[0]:.10f}",
              file=sys.stderr, flush=True)
        return out

    if args.copy_calibration_mb > 0:
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        nbytes = args.copy_calibration_mb << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        hip.hipMalloc(ctypes.byref(a), nbytes)
        hip.hipMalloc(ctypes.byref(b), nbytes)
        hip.hipMemcpy(b, a, nbytes, 3)           # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
        print(f"# copy calibration: {nbytes} bytes device to device", file=sys.stderr, flush=True)

    if args.forms == "adjoint":
        print(json.dumps({"partial_run": "adjoint", "adjoint": adjoint_leg(sph2d, args.adjoint_steps)}), flush=True)
        return 0
    if args.forms == "ms":
        print(json.dumps({"partial_run": "ms",
                          "moving_square": moving_square_leg(sph2d, args.ms_steps,
                                                             [float(r) for r in args.ms_re.split(",")])}),
              flush=True)
        return 0
    if args.forms == "batched":
        print(json.dumps({"partial_run": "batched",
                          "batched": batched_leg(sph2d, args.batch_instances, args.batch_steps, args.warmup,
                                                 args.batch_trial_steps)}),
              flush=True)
        return 0
    r64 = run_form(sph2d.FP64) if args.forms in ("both", "fp64") else None
    r32 = run_form(sph2d.FP32_MIXED) if args.forms in ("both", "mixed") else None
    if r64 is None or r32 is None:
        r = r64 or r32
        print(json.dumps({"partial_run": args.forms, "ms_per_step": 1000.0 * r["total_s"] / args.steps,
                          "force_pass_ms": r["force_ms"], "pass_a_ms": r["passa_ms"],
                          "evaluation_ms": r["eval_ms"]}), flush=True)
        return 0
    N = n * n
    sps64 = N * args.steps / r64["total_s"]
    sps32 = N * args.steps / r32["total_s"]
    ke_err = abs(r32["sample"][0] - r64["sample"][0]) / r64["sample"][0]
    t_end = (args.warmup + args.steps) * r64["dt"]
    # the dominant neighbour pass of the mixed form in this run, against the HBM roofline
    # with the frozen algorithmic bytes of that pass (PASS_BYTES)
    dom = "force" if r32["force_ms"] >= r32["passa_ms"] else "pass_a"
    dom_ms = r32["force_ms"] if dom == "force" else r32["passa_ms"]
    dom_kernel = "k_eval_c32 (FP32-mixed force pass)" if dom == "force" else "k_eval_a32 (FP32-mixed pass A)"
    alg32 = PASS_BYTES[dom][32] * r32["n_local"]
    ach32 = alg32 / (dom_ms * 1e-3) / 1e9
    traffic = PASS_TRAFFIC.get(dom)
    # the FP64 form's dominant pass (its force pass) on the same model
    ach64 = PASS_BYTES["force"][64] * r64["n_local"] / (r64["force_ms"] * 1e-3) / 1e9
    steps_per_s = args.steps / r32["total_s"]
    step_ach32 = STEP_BYTES[32] * sps32 / 1e9
    step_ach64 = STEP_BYTES[64] * sps64 / 1e9
    out = {
        "metric": "particle-steps/s, 4.2M-particle TGV, FP32-mixed vs FP64 at stated KE error",
        "value": sps32,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * r32["total_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed",
        "data": "synthetic (Taylor-Green lattice initial condition generated on the GPU)",
        "config": {
            "workload": f"Taylor-Green vortex {n}x{n} ({N} fluid + {r32['n_ghost']} ghost particles), "
                        f"predictor-corrector (1 force evaluation per step), modified linear Riemann "
                        f"+ delta-SPH + SPS, shifting, MLS every 20 steps; value = FP32-mixed form",
            "n_side": n, "dt": r32["dt"],
            "fp64_particle_steps_per_s": sps64,
            "fp64_ms_per_step": 1000.0 * r64["total_s"] / args.steps,
            "mixed_over_fp64_speedup": sps32 / sps64,
            "ke_rel_err_mixed_vs_fp64": ke_err,
            "ke_err_at_t": t_end,
            "ke_over_ke0_fp64": r64["sample"][0], "ke_over_ke0_mixed": r32["sample"][0],
            "pairs": r32["pairs"],
            "pair_interactions_per_s": r32["pairs"] * steps_per_s,
            "force_pass_ms": {"fp64": r64["force_ms"], "mixed": r32["force_ms"]},
            "pass_a_ms": {"fp64": r64["passa_ms"], "mixed": r32["passa_ms"]},
            "evaluation_ms": {"fp64": r64["eval_ms"], "mixed": r32["eval_ms"]},
            "fp64_force_pass_roofline": {"achieved": ach64, "unit": "GB/s", "frac": ach64 / HBM_PEAK_GBS,
                                         "bytes_per_particle": PASS_BYTES["force"][64]},
        },
        "roofline": {
            "kernel": dom_kernel,
            "bound": "hbm",
            "achieved": ach32,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach32 / HBM_PEAK_GBS,
            "traffic": (traffic["bytes_per_particle"] * r32["n_local"]
                        if traffic is not None and n == traffic["n_side"] else None),
            "traffic_source": traffic["source"] if traffic is not None else None,
            "bytes_per_particle": PASS_BYTES[dom][32],
            "bytes_model": "frozen algorithmic bytes of this pass per particle",
            "units_per_launch": r32["n_local"],
            "duration_ms": dom_ms,
            "launches_timed": r32["n_eval_timed"],
            # what the pass reads and writes for its own particle (records, flags, the
            # neighbour record, outputs), a per-launch count beside the frozen model
            "bytes_moved_per_particle": PASS_READ_WRITE[dom],
        },
        # the whole step on the same frozen model (one evaluation per predictor-corrector step)
        "step_roofline": {
            "bytes_per_particle_step": {"mixed": STEP_BYTES[32], "fp64": STEP_BYTES[64]},
            "achieved_gbs": {"mixed": step_ach32, "fp64": step_ach64},
            "frac": {"mixed": step_ach32 / HBM_PEAK_GBS, "fp64": step_ach64 / HBM_PEAK_GBS},
        },
    }
    if dom in PASS_VALU:
        flop = PASS_VALU[dom]["fp32_flop_per_particle"] * r32["n_local"] / (dom_ms * 1e-3) / 1e12
        out["valu_roofline"] = {
            "kernel": dom_kernel, "bound": "valu", "achieved": flop, "unit": "TFLOP/s",
            "peak": FP32_VECTOR_TFLOPS, "frac": flop / FP32_VECTOR_TFLOPS,
            "peak_unpacked": FP32_VECTOR_TFLOPS_UNPACKED, "frac_unpacked": flop / FP32_VECTOR_TFLOPS_UNPACKED,
            "note": "packed FP32 instructions are disabled in this build, so scalar FP32 issues at the FP64 "
                    "vector rate (78.6 TFLOP/s); 157.3 is the published (packed) FP32 peak",
            "fp64_achieved": PASS_VALU[dom]["fp64_flop_per_particle"] * r32["n_local"] / (dom_ms * 1e-3) / 1e12,
            "source": PASS_VALU[dom]["source"],
        }
    if args.ke_gate_side > 0:
        try:
            out["ke_gate"] = ke_gate_leg(sph2d, args.ke_gate_side)
        except Exception as e:                    # the headline line is printed regardless
            out["ke_gate"] = {"error": repr(e)}
    if args.batch_instances > 0:
        try:
            out["batched"] = batched_leg(sph2d, args.batch_instances, args.batch_steps, args.warmup,
                                         args.batch_trial_steps)
        except Exception as e:                    # the headline line is printed regardless
            out["batched"] = {"error": repr(e)}
    if args.adjoint_steps > 0:
        try:
            out["adjoint"] = adjoint_leg(sph2d, args.adjoint_steps)
        except Exception as e:                    # the headline line is printed regardless
            out["adjoint"] = {"error": repr(e)}
    if args.ms_steps > 0:
        try:
            out["moving_square"] = moving_square_leg(sph2d, args.ms_steps, [float(v) for v in args.ms_re.split(",")])
        except Exception as e:                    # the headline line is printed regardless
            out["moving_square"] = {"error": repr(e)}
    if args.adjoint_side > 0:
        try:
            out.setdefault("adjoint", {})["evaluation_vjp"] = evaluation_vjp_leg(sph2d, args.adjoint_side, 5)
        except Exception as e:                    # the headline line is printed regardless
            out.setdefault("adjoint", {})["evaluation_vjp"] = {"error": repr(e)}
    if rank == 0 and not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        head = cpu_baseline_all_cores(args.cpu_all_side, args.cpu_all_steps)
        samples = {}
        try:
            samples[f"{args.cpu_side}_one_core"] = cpu_baseline(args.cpu_side, args.cpu_steps)
            print("# cpu baseline: small lattice on all cores ...", file=sys.stderr, flush=True)
            samples[f"{args.cpu_small_side}_all_cores"] = cpu_baseline_all_cores(args.cpu_small_side,
                                                                               args.cpu_small_steps)
        except Exception as e:                    # the headline line is printed regardless
            samples["error"] = repr(e)
        out["cpu_baseline"] = dict(head, samples=samples)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
