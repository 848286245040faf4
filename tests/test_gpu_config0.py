"""The reference-runnable Taylor-Green case as its configuration states it: 100^2 lattice,
rho0 = 1, U0 = 1, Re = 100, c0 = 10 U0, predictor-corrector at the CFL time step of
Initialize_SPH_Calculation.f90:449-461 with the 0.2 coefficient
(2 x 0.1 x min(Morris 0.125 h^2 / 1.01e-6, Monaghan 0.25 h / (1.1 c0)) = 5.4545e-5,
REAL(4) literals), to t = 0.5.

* 25 steps against the oracle at the parity bars;
* 9,167 steps to t = 0.5 against the oracle's run of the same case, sampled every
  1,000 steps (to rounding until step 1,000, then the integral quantities), and
  the kinetic energy at t = 0.5 against the analytic decay.

rho0 = 1 is not the rho0 = 1000 run rescaled: the SPS term's carried temp_p
(Solver...SPS.f90:458-485) scales as 1/rho0^2 and is a few permille of the
acceleration at rho0 = 1 (pinned on the oracle by the density-scaling test).
"""
import math
import os

import numpy as np
import pytest

import sph2d
from oracle.pyoracle import Oracle
from test_gpu_parity import rel_err, say

pytestmark = pytest.mark.gpu

N = 100


def cfl_dt(h, c0):
    f32 = lambda v: float(np.float32(v))
    morris = 0.125 * h * h / f32(1.01e-6)
    monaghan = 0.25 * h / (f32(1.1) * math.sqrt(c0 * c0))
    return float(np.float32(2) * np.float32(0.1)) * min(morris, monaghan)


def params(rho0=1.0):
    dx = 1.0 / N
    h = 1.2 * dx
    return sph2d.tgv_params(N, integrator=sph2d.INTEGRATOR_PC, rho0=rho0, dt=cfl_dt(h, 10.0))


def test_config0_dt():
    prm = params()
    assert abs(prm.dt - 5.4545e-5) < 1e-8


def test_config0_predictor_corrector_against_the_oracle():
    say("test_config0_predictor_corrector_against_the_oracle")
    prm = params()
    sim = sph2d.Simulation(prm)
    o = Oracle(nx=N, ny=N, integrator=2, dt=prm.dt, rho0=1.0)
    o.init_tgv()
    st = tuple(o.f64(f)[:N * N].copy() for f in ("g_x", "g_y", "g_u", "g_v", "g_rho", "g_mass"))
    sim.upload_state(*st)
    for it in range(1, 26):
        sim.time_step(it)
        o.time_step(it)
    g = sim.download_state()
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v"), ("rho", "g_rho"), ("p", "g_p")):
        r = o.f64(of)[:N * N]
        e = float(np.max(np.abs(g[f] - r))) if f in ("x", "y") else rel_err(g[f], r)
        assert e < (1e-12 if f in ("x", "y") else 1e-10), (f, e)
    sim.close()
    o.close()


def test_config0_run_to_t05_against_the_oracle_run():
    """9,167 steps against the oracle's run of the same case (tests/golden/make_config0.py
    keeps its samples every 1,000 steps).

    Two FP64 implementations whose sums round differently agree to rounding only
    until the particle arrangement starts to depend on discrete events: a lattice
    pair crossing the 3h cutoff (where the truncated kernel's gradient is not zero)
    one step earlier in one run, a shifting clamp, a surface flag. After ~1,500 steps
    the two runs' particle-level fields part (the error norms' maxima, single
    particles, first); the integral quantities do not. Measured on the MI355X: at step
    1,000 KE/KE0 agrees to 5e-14 and the error norms to 2e-7; after that KE/KE0 stays
    within 1e-4 (9.7e-5 at t = 0.5), max|v| within 4e-3, and the L1/L2 error norms,
    themselves of order 1e-3, within 15%. The bars: step 1,000 at 1e-10 (KE, max|v|)
    and 1e-5 (norms); every row KE/KE0 within 2e-4, max|v| within 1e-2 and the L2
    norms within 30%; KE/KE0 at t = 0.5 within 6% of the analytic decay
    exp(-16 pi^2 nu t) (5.0% below it here; the reference's own 100^2 run ends 5%
    below it, Kinetic_energy.dat)."""
    say("test_config0_run_to_t05_against_the_oracle_run")
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "tgv100_pc_rho1.npz"))
    rows = gold["samples"]
    prm = params()
    assert prm.dt == float(gold["dt"])
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    done = 0
    for row in rows:
        it = int(row[0])
        sim.run(done + 1, it - done)
        done = it
        s, r = sim.sample(), row[2:]
        rel = np.abs(s - r) / np.abs(r)
        say(it, f"KE/KE0 {s[0]:.10f} oracle {r[0]:.10f}; relative differences: KE {rel[0]:.2e} max|v| {rel[1]:.2e} "
                f"L2 norms {rel[[4, 7, 10, 13]].max():.2e}")
        if it <= 1000:
            assert rel[0] < 1e-10 and rel[1] < 1e-10 and rel[2:].max() < 1e-5, (it, rel)
        assert rel[0] < 2e-4 and rel[1] < 1e-2 and rel[[4, 7, 10, 13]].max() < 0.3, (it, rel)
    t = done * prm.dt
    ka = math.exp(-16 * math.pi ** 2 * prm.nu * t)
    assert abs(t - 0.5) < prm.dt and abs(s[0] - ka) / ka < 0.06, (s[0], ka)
    sim.close()
