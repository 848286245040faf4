"""GPU parity of the moving-square program (ms = 1; the reference's
Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/) against the oracle, and the
moving-square workload at dx = D/100 through property checks.

Bars as in test_gpu_parity: local order bit-exact; the FP64 rates per particle within
1e-12 of max(|ref|, 1e-3 sum |terms|); fixed-ghost state and intermediate fields within
1e-12 of the field's largest magnitude; 21 predictor-corrector steps within 1e-10
(positions 1e-12 absolute); the FP32-mixed form within 1e-4 of the field norm.
States are perturbed with seeded noise so that no MUSCL slope is exactly zero (the
limiter is discontinuous there and the last bit of a gather sum would pick the branch).
"""
import sys

import numpy as np
import pytest

import sph2d
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

TANK = dict(nx=100, ny=48, box_x=2.0, box_y=0.96, body_cx=0.6, body_cy=0.48, body_len=0.4)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(g, r):
    scale = np.max(np.abs(r))
    return float(np.max(np.abs(g - r)) / scale) if scale > 0 else float(np.max(np.abs(g)))


def particle_rel_err(g, r, mag):
    den = np.maximum(np.abs(r), 1e-3 * mag)
    q = np.where(den > 0, np.abs(g - r) / np.where(den > 0, den, 1.0), np.abs(g - r))
    return float(np.max(q)) if len(q) else 0.0


def params(solver=sph2d.SOLVER_MS_MUSCL_HLLC, sps_on=1, precision=sph2d.FP64, **kw):
    p = sph2d.moving_square_params(n_per_d=50, re=50.0, tank=(2.0, 0.96), layers=3, solver=solver, sps_on=sps_on,
                                   precision=precision, **TANK)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def oracle_for(p):
    o = Oracle(ms=1, nx=p.nx, ny=p.ny, box_x=p.box_x, box_y=p.box_y, body_cx=p.body_cx, body_cy=p.body_cy,
               body_len=p.body_len, walls=p.walls, rho0=p.rho0, c0sq=p.c0sq, nu=p.nu, h_factor=1.15, dt=p.dt,
               integrator=2, solver=p.solver, sps_on=p.sps_on, shift_on=p.shift_on, mls_on=p.mls_on,
               mls_every=p.mls_every, mls_order=3, mls_alpha=p.mls_alpha, alpha=p.alpha, beta=p.beta, fai=p.fai,
               delta=p.delta, delta_alpha=p.delta_alpha, epsilon_xsph=p.epsilon_xsph, background_p=p.background_p,
               release_time=p.release_time, body_speed=p.body_speed, body_ramp=p.body_ramp,
               sample_every=p.sample_every, max_shift_dx=0.15, gravity=p.gravity)
    o.init_moving_square()
    return o


def noisy_pair(p, seed=7):
    """GPU context and oracle on the same perturbed tank state."""
    sim = sph2d.Simulation(p, capacity=p.nx * p.ny)
    sim.init_moving_square()
    o = oracle_for(p)
    n = o.n_fluid
    assert sim.n_fluid == n
    rng = np.random.default_rng(seed)
    x, y = o.f64("g_x")[:n].copy(), o.f64("g_y")[:n].copy()
    u = 0.2 * np.sin(3 * x) * np.cos(5 * y) + 0.01 * rng.standard_normal(n)
    v = 0.2 * np.cos(2 * x) * np.sin(4 * y) + 0.01 * rng.standard_normal(n)
    rho = p.rho0 * (1.0 + 2e-3 * np.cos(4 * x + y) + 2e-4 * rng.standard_normal(n))
    m = o.f64("g_mass")[:n].copy()
    for nm, a in (("g_u", u), ("g_v", v), ("g_rho", rho)):
        g = o.f64(nm)
        g[:n] = a
        o.set_f64(nm, g)
    sim.upload_state(x, y, u, v, rho, m)
    return sim, o


SOLVERS = [(sph2d.SOLVER_MS_MUSCL_HLLC, 1), (sph2d.SOLVER_MS_HLLC, 1), (sph2d.SOLVER_MS_HLLC, 0),
           (sph2d.SOLVER_MODIFIED_LINEAR_RIEMANN, 0)]


@pytest.mark.parametrize("solver,sps_on", SOLVERS)
def test_ms_single_evaluation_parity(solver, sps_on):
    say("test_ms_single_evaluation_parity", solver, sps_on)
    p = params(solver=solver, sps_on=sps_on, sample_every=1)
    sim, o = noisy_pair(p)
    t = 0.2
    sim.body_motion(t)
    o.ms_body_motion(t)
    sim.devide_block(1)
    o.devide_block(1)
    sim.single_step_compute(1)
    o.ms_single_step_compute(1, t)
    assert np.array_equal(sim.local("name"), o.i32("name"))
    typ = o.i32("type")
    assert np.array_equal(sim.local("type"), typ)
    w = typ == -3
    for f in ("p", "u", "v", "rho", "mass"):
        e = rel_err(sim.local(f)[w], o.f64(f)[w])
        say("ghost", f, e)
        assert e < 1e-12, f
    for g, r in (("grad_rho_x", "grad_rho_x"), ("grad_rho_y", "grad_rho_y"), ("avg_x", "avg_vel_x"),
                 ("avg_y", "avg_vel_y")):
        e = rel_err(sim.local(g), o.f64(r))
        say(g, e)
        assert e < 1e-12, g
    f = typ == 2
    for k in ("drhodt", "ax", "ay"):
        e = particle_rel_err(sim.local(k)[f], o.f64(k)[f], o.f64("mag_" + k)[f])
        say(k, e)
        assert e < 1e-12, k
    o.shift_sums()                 # Shift_particle_position.f90:78-103 over this evaluation's pairs
    for k in ("shift_x", "shift_y"):
        e = rel_err(sim.local(k)[f], o.f64(k)[f])
        say(k, e)
        assert e < 1e-12, k
    bf, step = sim.body_force()
    ob = o.ms_body_force()
    say("body force", bf, ob)
    assert step == 1
    assert np.all(np.abs(bf - ob) <= 1e-12 * np.abs(ob[[0, 1, 2, 3, 0, 1]]).max() + 1e-12 * np.abs(ob).max())
    sim.close()


@pytest.mark.parametrize("solver,sps_on", [(sph2d.SOLVER_MS_MUSCL_HLLC, 1), (sph2d.SOLVER_MODIFIED_LINEAR_RIEMANN, 0)])
def test_ms_steps_parity(solver, sps_on):
    """21 MS predictor-corrector steps (body motion, ghosts, the MLS of step 20, shifting,
    the body force of steps 10 and 20) against the oracle."""
    say("test_ms_steps_parity", solver, sps_on)
    p = params(solver=solver, sps_on=sps_on, mls_every=20, sample_every=10)
    sim, o = noisy_pair(p, seed=11)
    for it in range(1, 22):
        sim.time_step(it)
        o.ms_time_step(it)
        if it == 20:
            bf, step = sim.body_force()
            ob = o.ms_body_force()
            say("body force at 20", bf, ob)
            assert step == 20 and rel_err(bf, ob) < 1e-10
    g = sim.download_state()
    n = o.n_fluid
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v"), ("rho", "g_rho")):
        r = o.f64(of)[:n]
        e = float(np.max(np.abs(g[f] - r))) if f in ("x", "y") else rel_err(g[f], r)
        say(f, e)
        assert e < (1e-12 if f in ("x", "y") else 1e-10), (f, e)
    sim.close()


def test_ms_mixed_form_against_the_fp64_oracle():
    """The FP32-mixed MS evaluation (FP32 pair sums with compensation, FP64 positions, EOS,
    ghosts and stages) against the FP64 oracle: every rate and the body force within 1e-4
    of the field's norm."""
    say("test_ms_mixed_form_against_the_fp64_oracle")
    p = params(precision=sph2d.FP32_MIXED, sample_every=1)
    sim, o = noisy_pair(p)
    t = 0.2
    sim.body_motion(t)
    o.ms_body_motion(t)
    sim.devide_block(1)
    o.devide_block(1)
    sim.single_step_compute(1)
    o.ms_single_step_compute(1, t)
    f = o.i32("type") == 2
    for k in ("drhodt", "ax", "ay"):
        e = rel_err(sim.local(k)[f], o.f64(k)[f])
        say("mixed", k, e)
        assert e < 1e-4, k
    bf, _ = sim.body_force()
    ob = o.ms_body_force()
    e = float(np.max(np.abs(bf - ob)) / np.abs(ob).max())
    say("mixed body force", bf, ob, e)
    assert e < 1e-4
    sim.close()


def test_ms_moving_square_at_d_over_100():
    """The moving-square workload (10 D x 5 D tank, dx = D/100: 490,000 fluid particles,
    3 ghost layers, Re = 100, MUSCL-HLLC + delta-SPH + SPS + shifting) through property
    checks: 200 steps in both forms, the body moved by its closed form, every state
    finite, the force on the water finite with lift far below drag, and the two forms'
    drag within 1%."""
    say("test_ms_moving_square_at_d_over_100")
    forces = {}
    for prec in (sph2d.FP64, sph2d.FP32_MIXED):
        p = sph2d.moving_square_params(n_per_d=100, re=100.0, precision=prec, sample_every=100)
        sim = sph2d.Simulation(p, capacity=p.nx * p.ny)
        sim.init_moving_square()
        assert sim.n_fluid == 490000
        for it in range(1, 201):
            sim.time_step(it)
        g = sim.download_state()
        for k in ("x", "y", "u", "v", "rho"):
            assert np.all(np.isfinite(g[k])), k
        assert np.all(g["x"] > 0) and np.all(g["x"] < 10) and np.all(g["y"] > 0) and np.all(g["y"] < 5)
        bf, step = sim.body_force()
        say("precision", prec, "body force at", step, bf)
        assert step == 200 and np.all(np.isfinite(bf))
        assert abs(bf[5]) < 1e-2 * abs(bf[4])
        forces[prec] = bf
        sim.close()
    e = abs(forces[sph2d.FP32_MIXED][4] - forces[sph2d.FP64][4]) / abs(forces[sph2d.FP64][4])
    say("mixed vs FP64 drag", e)
    assert e < 1e-2
