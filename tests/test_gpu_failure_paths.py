"""Device-side failure paths of the step and the 2x2 renormalisation mode, against the
oracle:

* ill particles: fluid below y = 0 and fluid outside the background grid are typed
  120 (Devide_block_and_send_tasks.f90:75-83,141-155), leave the pair sums and the
  stages, n_ill counts them, and the refresh leaves zero entries of type 0; the rest
  of the tank stays within the parity bars;
* the periodic ghost capacity: a state whose ghosts outnumber the context's ghost
  capacity (all particles inside the left edge's 3h band) makes the emission clamp
  its writes and flag the overflow, reported as SPH2D_ERR_CAPACITY;
* a non-finite state is refused with SPH2D_ERR_NONFINITE;
* renormalization = 1 (the dyadic tensor's inverse instead of the identity) on a
  holed disc with a free surface: per-particle rates of one evaluation at the parity
  bar and 21 predictor-corrector steps, and n_singular equal to the oracle's on a
  cloud of isolated pairs and particles.
"""
import numpy as np
import pytest

import sph2d
from cases import hydrostatic_tank
from oracle.pyoracle import Oracle
from test_gpu_parity import assert_rates, evaluate_both, rel_err, say
from test_oracle_pins import holed_disc

pytestmark = pytest.mark.gpu


def test_ill_particles_are_typed_120_as_the_oracle_does():
    """Typed 120 by the step's binning, left out of the subdomain, then zero entries of
    type 0 after the refresh (Refresh_domain rebuilds the global arrays from the
    subdomain), skipped by every later binning."""
    say("test_ill_particles_are_typed_120_as_the_oracle_does")
    n, H, g = 40, 0.5, 9.81
    x, y, u, v, rho, m = [a.copy() for a in hydrostatic_tank(n, H, g)]
    M = len(x)
    y[5] = -0.004                                   # below the bed, inside the walls
    x[7] = 5.0                                      # beyond the background grid
    st = (x, y, u, v, rho, m)
    kw = dict(integrator=2, walls=4, gravity=g)
    prm = sph2d.tgv_params(n, periodic=0, dt=1e-4 * 100 / n * 0.5, **kw)
    sim = sph2d.Simulation(prm, capacity=n * n)
    sim.upload_state(*st)
    o = Oracle(nx=n, ny=n, dt=prm.dt, **kw)
    o.set_fluid_state(*st)
    for it in range(1, 6):
        sim.time_step(it)
        o.time_step(it)
        if it == 1:
            assert sim.diagnostics().n_ill == 2
    assert sim.diagnostics().n_ill == 0             # type 0 from the first refresh on
    gs = sim.download_state()
    t = o.i32("g_type")[:M]
    assert t[5] == 0 and t[7] == 0 and np.count_nonzero(t != 2) == 2
    assert np.array_equal(gs["type"], t)
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v"), ("rho", "g_rho"), ("mass", "g_mass")):
        r = o.f64(of)[:M]
        e = float(np.max(np.abs(gs[f] - r))) if f in ("x", "y") else rel_err(gs[f], r)
        assert e < (1e-12 if f in ("x", "y") else 1e-10), (f, e)
        assert gs[f][5] == 0.0 and gs[f][7] == 0.0, f
    s1, s2 = sim.sample(), o.sample()                       # the samples leave them out too
    assert abs(s1[1] - s2[1]) < 1e-10 * abs(s2[1])          # max |v| (KE/KE0: the tank starts at rest)
    sim.close()
    o.close()


def test_ghost_capacity_overflow_is_reported():
    """A 100^2 box sizes its ghost capacity for lattice-like states (16,384 here);
    20,000 particles inside the left edge's 3h band each need an x-ghost."""
    say("test_ghost_capacity_overflow_is_reported")
    n = 100
    prm = sph2d.tgv_params(n, integrator=2)
    N = 20000
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 2 * prm.h, N)
    y = rng.uniform(0.0, 1.0, N)
    rho = np.full(N, 1000.0)
    sim = sph2d.Simulation(prm, capacity=N)
    sim.upload_state(x, y, np.zeros(N), np.zeros(N), rho, rho * 1e-4)
    sim.distribute_period_boundary()
    with pytest.raises(sph2d.SphError, match="ghost capacity") as ei:
        sim.diagnostics()
    assert ei.value.code == sph2d.ERR_CAPACITY
    sim.close()


def test_non_finite_state_is_refused():
    say("test_non_finite_state_is_refused")
    n = 20
    prm = sph2d.tgv_params(n, integrator=2)
    sim = sph2d.Simulation(prm)
    N = n * n
    x = (np.arange(N) % n + 0.5) / n
    y = (np.arange(N) // n + 0.5) / n
    rho = np.full(N, 1000.0)
    u = np.zeros(N)
    u[3] = np.nan
    with pytest.raises(sph2d.SphError) as ei:
        sim.upload_state(x, y, u, np.zeros(N), rho, rho / N)
    assert ei.value.code == sph2d.ERR_NONFINITE
    rho[9] = -1.0
    with pytest.raises(sph2d.SphError) as ei:
        sim.upload_state(x, y, np.zeros(N), np.zeros(N), rho, np.full(N, 1.0 / N))
    assert ei.value.code == sph2d.ERR_NONFINITE
    sim.close()


def renorm_pair(n, st):
    prm = sph2d.tgv_params(n, integrator=2, renormalization=1)
    sim = sph2d.Simulation(prm)
    sim.upload_state(*st)
    o = Oracle(nx=n, ny=n, integrator=2, dt=prm.dt, renorm=1)
    o.set_fluid_state(*st)
    return sim, o


def test_renormalised_evaluation_and_steps_against_the_oracle():
    say("test_renormalised_evaluation_and_steps_against_the_oracle")
    n = 60
    st = holed_disc(n)
    sim, o = renorm_pair(n, st)
    evaluate_both(sim, o)
    assert sim.diagnostics().n_surface > 20
    assert_rates(sim, o)
    for f in ("grad_rho_x", "grad_rho_y", "tau_xx", "normal_x", "normal_y"):
        assert rel_err(sim.local(f), o.f64(f)) < 1e-12, f
    assert np.array_equal(sim.local("surface"), o.i32("surface"))
    sim.close()
    sim, o = renorm_pair(n, st)
    for it in range(1, 22):
        sim.time_step(it)
        o.time_step(it)
    gs = sim.download_state()
    M = len(st[0])
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v"), ("rho", "g_rho")):
        r = o.f64(of)[:M]
        e = float(np.max(np.abs(gs[f] - r))) if f in ("x", "y") else rel_err(gs[f], r)
        say(f, e)
        assert e < (1e-12 if f in ("x", "y") else 1e-10), (f, e)
    # the mode changes the flow near the surface by far more than the bar
    plain = sph2d.Simulation(sph2d.tgv_params(n, integrator=2))
    plain.upload_state(*st)
    for it in range(1, 22):
        plain.time_step(it)
    assert rel_err(plain.download_state()["u"], gs["u"]) > 1e-6
    for s in (sim, plain):
        s.close()
    o.close()


def test_singular_tensors_counted_as_the_oracle_counts():
    say("test_singular_tensors_counted_as_the_oracle_counts")
    n = 40
    dx = 1.0 / n
    pts = []
    for k in range(6):
        x0 = 0.1 + 0.13 * k
        pts += [(x0, 0.2), (x0 + dx, 0.2)]
    pts += [(0.2, 0.6), (0.5, 0.6), (0.8, 0.6)]
    for i in range(5):
        for j in range(5):
            pts.append((0.4 + i * dx, 0.85 + j * dx * 0.5))
    p = np.array(pts)
    k = len(p)
    rho = np.full(k, 1000.0)
    st = (p[:, 0], p[:, 1], np.zeros(k), np.zeros(k), rho, rho * dx * dx)
    sim, o = renorm_pair(n, st)
    evaluate_both(sim, o)
    assert o.n_singular == 15
    assert sim.diagnostics().n_singular == o.n_singular
    assert_rates(sim, o)
    sim.close()
    o.close()
