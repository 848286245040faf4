"""Throughput of the WCSPH step on one MI355X: Taylor-Green vortex on a 2048^2
lattice (4,194,304 fluid particles + 32,832 periodic ghosts), predictor-corrector
(time_iteration.f90, leapforg_or_eular=2), plain FP64 form.

One step = one pass of time_iteration.f90's loop body: periodic ghosts, cell-key
sort + reorder, stage updates with one force evaluation (two on step 1), MLS
density reinitialisation every 20 steps, particle shifting and the refresh.
The state is resident in HBM when the timed region starts; timing uses HIP
events on the context's stream. Prints ONE JSON line.

Also reported: the force pass's roofline (algorithmic bytes per launch over its
HIP-event-timed duration) and the CPU baseline (the C oracle, one core, on a
bounded sample of the same workload).
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


def force_pass_bytes_per_particle() -> int:
    """Algorithmic HBM bytes per subdomain particle of one force-pass launch (k_eval_c):
    every array the pass reads for its own particle once, every output once; the
    neighbours' data are the same arrays (served from L2/L1), not counted again.
    reads:  pos 16, vel 16, (rho,V) 16, (p,c) 16, type 4, cell 4, mass 8,
            global mass at local index 8, tau 24, grad rho 16, near flag 4 = 132
    writes: drho/dt 8, acceleration 16, shift 16 = 40."""
    return 132 + 40


def cpu_baseline(n_side: int = 512, steps: int = 4) -> dict:
    """The C oracle (oracle/liboracle.so, the reference's algorithm restated,
    scatter over a half pair list as in find_particle_pair.f90), one core, on a
    bounded sample: n_side^2 TGV, predictor-corrector, `steps` steps."""
    from oracle.pyoracle import Oracle
    o = Oracle(nx=n_side, ny=n_side, integrator=2, dt=1e-4 * 100.0 / n_side)
    o.init_tgv()
    o.time_step(1)                                # step 1 has the extra evaluation
    t0 = time.perf_counter()
    for it in range(2, 2 + steps):
        o.time_step(it)
    dt = time.perf_counter() - t0
    o.close()
    n = n_side * n_side
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"TGV {n_side}^2 ({n} particles), predictor-corrector FP64, steps 2..{1 + steps} "
                      f"timed ({dt:.1f} s); oracle/liboracle.so single-threaded"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n-side", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-side", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every GPU box here has one, and the "
                                   "multi-GPU form (spatial slabs with a halo exchange per evaluation) "
                                   "is designed but not built"}))
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import sph2d
    n = args.n_side
    prm = sph2d.tgv_params(n, integrator=sph2d.INTEGRATOR_PC)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    print(f"# setup {n}^2 done", file=sys.stderr, flush=True)
    if args.warmup > 0:
        sim.run(1, args.warmup)
    d = sim.diagnostics()                         # synchronises
    first = 1 + args.warmup
    hip = hip_events()
    stream = ctypes.c_void_p(sim.stream())
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipEventRecord(e0, stream)
    sim.run(first, args.steps)
    hip.hipEventRecord(e1, stream)
    hip.hipEventSynchronize(e1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    total_s = ms.value / 1000.0
    s = sim.sample()
    d = sim.diagnostics()
    if d.error_flags or d.rebin_needed:
        print(f"# warning: error_flags={d.error_flags} rebin={d.rebin_needed}", file=sys.stderr)
    print(f"# timed {args.steps} steps in {total_s:.3f} s; KE/KE0={s[0]:.9f}", file=sys.stderr, flush=True)

    # roofline of the dominant kernel: the force pass, timed alone with HIP events
    reps = 20
    force_ms = sim.time_pass(0, reps)
    passa_ms = sim.time_pass(1, reps)
    eval_ms = sim.time_pass(2, reps)
    n_local = d.n_local
    bytes_launch = force_pass_bytes_per_particle() * n_local
    achieved = bytes_launch / (force_ms * 1e-3) / 1e9
    # pair count of the state (for pair interactions/s)
    cnt = ctypes.c_int64()
    sph2d._check(sph2d.lib().sph2d_download_pairs(sim._c, 0, None, None, ctypes.byref(cnt)))
    pairs = cnt.value
    steps_per_s = args.steps / total_s
    value = n * n * steps_per_s
    evals_per_step = 1
    out = {
        "metric": "particle-steps/s, 4.2M-particle TGV, FP32-mixed vs FP64 at stated KE error",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * total_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green lattice initial condition generated on the GPU)",
        "config": {
            "workload": f"Taylor-Green vortex {n}x{n} ({n * n} fluid + {d.n_ghost} ghost particles), "
                        f"predictor-corrector (1 force evaluation per step), modified linear Riemann "
                        f"+ delta-SPH + SPS, shifting, MLS every 20 steps, plain FP64 form",
            "n_side": n, "dt": prm.dt, "form": "fp64",
            "ke_over_ke0_at_end": s[0],
            "pairs": pairs,
            "pair_interactions_per_s": pairs * evals_per_step * steps_per_s,
            "force_pass_ms": force_ms, "pass_a_ms": passa_ms, "evaluation_ms": eval_ms,
        },
        "roofline": {
            "kernel": "k_eval_c (force pass)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_particle": force_pass_bytes_per_particle(),
            "units_per_launch": n_local,
        },
    }
    if rank == 0 and not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_side, args.cpu_steps)
    sim.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
