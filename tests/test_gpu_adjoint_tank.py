"""The differentiable wave tank: sph2d_forward / sph2d_backward.

A walled tank (the unit box, still water 0.5 deep on a 40^2 lattice, jittered by
+-0.03 dx so that no two distances, angles or heights tie) with three columns of
piston wavemaker particles, gravity and a linear end damping zone; the paddle
stroke is x_p(t) = sum_k a_k sin(2 pi k f0 t) over 8 harmonics; the loss weights the
MLS pressure of three probes 0.1 below the still surface, sampled every 5 steps.

The loss is piecewise smooth: the branches the reverse holds fixed (neighbour sets,
surface flags, the water height's arg-max, ...) switch somewhere along almost any
line through parameter space, about once per 1e-6 of |a| here, each time with a
jump of ~1e-5 in the loss. Central differences are therefore taken with steps of
1e-8 and below (where the loss is smooth to its rounding, ~3e-10) and each
direction is judged by the best of three steps: a jump inside one step's stencil
shows as a difference ~ jump / (2 eps) that the smaller steps do not share.
"""
import math

import numpy as np
import pytest

import sph2d
from cases import hydrostatic_tank
from oracle.pyoracle import Oracle
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

N_SIDE, DEPTH, G = 40, 0.5, 9.81
PROBES = [(0.3, 0.4), (0.5, 0.4), (0.7, 0.4)]
KW = dict(integrator=2, walls=4, gravity=G, wavemaker_layers=3, damping=sph2d.DAMPING_LINEAR,
          damping_length=0.3, damping_amplitude=0.5, water_depth=DEPTH, mls_on=0, periodic=0)
EPS = (1e-8, 2e-9, 5e-10)
TOL = 1e-4


def tank_state():
    st = list(hydrostatic_tank(N_SIDE, DEPTH, G))
    jr = np.random.default_rng(11)
    st[0] = st[0] + jr.uniform(-0.03, 0.03, len(st[0])) / N_SIDE
    st[1] = st[1] + jr.uniform(-0.03, 0.03, len(st[1])) / N_SIDE
    return tuple(st)


def tank_params(**kw):
    return sph2d.tgv_params(N_SIDE, dt=1e-4 * 100 / N_SIDE * 0.5, **dict(KW, **kw))


def cfg(k=20, sample_every=5):
    return sph2d.adjoint_cfg(n_harmonics=8, f0=5.0, standing_steps=1, sample_every=sample_every, probes=PROBES,
                             checkpoint_every=k)


def make_sim(st, first=2, **kw):
    """A context at the start of the window: step 1 (and any steps before `first`)
    run with the paddle at rest."""
    sim = sph2d.Simulation(tank_params(**kw), capacity=N_SIDE * N_SIDE)
    sim.upload_state(*st)
    for it in range(1, first):
        sim.wavemaker_update(0.0, 0.0)
        sim.time_step(it)
    return sim


def paddle(a, f0, dt, it, standing=1):
    """The stroke of sph2d_forward, summed in its order."""
    if it <= standing:
        return 0.0, 0.0
    t = dt * (it - standing)
    xp = up = 0.0
    for k in range(1, len(a) + 1):
        om = 2.0 * math.pi * k * f0
        xp += float(a[k - 1]) * math.sin(om * t)
        up += float(a[k - 1]) * om * math.cos(om * t)
    return xp, up


@pytest.fixture(scope="module")
def setup():
    st = tank_state()
    rng = np.random.default_rng(5)
    a = rng.normal(0, 1e-3, 8)
    return st, a, rng


def loss(st, a, w, steps, k=20):
    sim = make_sim(st)
    p = sim.forward(cfg(k), a, 2, steps)
    sim.close()
    return float(np.sum(w * p))


def fd_best(f, a, d, g):
    """The central difference (over EPS) closest to g . d, and its relative error."""
    best = None
    for eps in EPS:
        fd = (f(a + eps * d) - f(a - eps * d)) / (2 * eps)
        e = abs(fd - g @ d) / max(abs(g @ d), 1e-30)
        if best is None or e < best[1]:
            best = (fd, e, eps)
    return best


def test_gradient_matches_central_differences_on_16_directions(setup):
    say("test_gradient_matches_central_differences_on_16_directions")
    st, a, rng = setup
    steps = 200
    sim = make_sim(st)
    p = sim.forward(cfg(), a, 2, steps)
    assert p.shape == (steps // 5, 3) and np.all(np.isfinite(p))
    w = rng.normal(size=p.shape)
    g = sim.backward(w)
    d0 = sim.diagnostics()
    assert d0.error_flags == 0 and d0.rebin_needed == 0
    sim.close()
    assert np.all(np.isfinite(g)) and np.max(np.abs(g)) > 0
    errs = []
    for j in range(16):
        d = rng.normal(size=8)
        d /= np.linalg.norm(d)
        fd, e, eps = fd_best(lambda aa: loss(st, aa, w, steps), a, d, g)
        say(f"direction {j}: fd {fd:.10e} adjoint {g @ d:.10e} rel {e:.2e} (eps {eps})")
        errs.append(e)
    assert max(errs) < TOL, errs


def test_checkpoint_interval_leaves_forward_and_gradient_unchanged(setup):
    """K = 1, 7 and 20 (a partial last segment) give the same probe pressures and the
    same gradient bit for bit; the backward leaves the forward's final state."""
    say("test_checkpoint_interval_leaves_forward_and_gradient_unchanged")
    st, a, rng = setup
    steps = 60
    out = {}
    for k in (20, 7, 1):
        sim = make_sim(st)
        p = sim.forward(cfg(k), a, 2, steps)
        before = sim.download_state()
        g = sim.backward(np.ones_like(p))
        after = sim.download_state()
        for f in ("x", "y", "u", "v", "rho", "p"):
            assert np.array_equal(before[f], after[f]), (k, f)
        out[k] = (p, g)
        sim.close()
    for k in (7, 1):
        assert np.array_equal(out[k][0], out[20][0]) and np.array_equal(out[k][1], out[20][1]), k


def test_forward_equals_the_step_loop_and_its_probes(setup):
    """sph2d_forward is sph2d_wavemaker_update + sph2d_time_step per step with the
    stroke's (x_p, u_p), and its probe rows are sph2d_sampling's MLS pressures."""
    say("test_forward_equals_the_step_loop_and_its_probes")
    st, a, _ = setup
    first, steps = 6, 30                       # samples after steps 10, 15, ... (sample_every 5)
    c = cfg(sample_every=5)
    sim = make_sim(st, first=first, sample_every=5)
    p = sim.forward(c, a, first, steps)
    end = sim.download_state()
    sim.close()
    sim = make_sim(st, first=first, sample_every=5)
    rows = []
    for k in range(steps):
        it = first + k
        xp, up = paddle(a, c.f0, sim.params.dt, it)
        sim.wavemaker_update(xp, up)
        sim.time_step(it)
        if (k + 1) % 5 == 0:
            assert it % 5 == 0
            rows.append(sim.sampling(it, [q[0] for q in PROBES], [q[1] for q in PROBES], G)["mls_p"])
    ref = sim.download_state()
    sim.close()
    assert np.array_equal(p, np.array(rows))
    for f in ("x", "y", "u", "v", "rho", "p"):
        assert np.array_equal(end[f], ref[f]), f


def test_autograd_function_of_a_misfit(setup):
    """torch.autograd.Function semantics: loss = mean((p - target)^2) with a target
    from other coefficients; loss.backward() equals sph2d_backward with the weights
    2 (p - target) / n."""
    import torch
    say("test_autograd_function_of_a_misfit")
    st, a, rng = setup
    steps = 60
    c = cfg()
    sim = make_sim(st)
    target = sim.forward(c, rng.normal(0, 1e-3, 8), 2, steps)
    sim.close()
    Fn = sph2d.paddle_probe_function()
    sims = []

    def fresh():
        s = make_sim(st)
        sims.append(s)
        return s

    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    p = Fn.apply(at, fresh, c, 2, steps)
    lossv = torch.mean((p - torch.from_numpy(target)) ** 2)
    lossv.backward()
    sim = make_sim(st)
    p2 = sim.forward(c, a, 2, steps)
    g2 = sim.backward(2 * (p2 - target) / p2.size)
    sim.close()
    for s in sims:
        s.close()
    assert np.array_equal(p.detach().numpy(), p2)
    np.testing.assert_allclose(at.grad.numpy(), g2, rtol=1e-12, atol=0)


def test_gradient_against_central_differences_of_the_oracle(setup):
    """An FP64 reference gradient: central differences of the same loss computed by
    the C oracle (its own wavemaker update, step and probe MLS) along 3 directions
    agree with the GPU adjoint to 1e-4."""
    say("test_gradient_against_central_differences_of_the_oracle")
    st, a, rng = setup
    steps = 200                                # long enough for the probes to answer the paddle
    c = cfg()
    sim = make_sim(st)
    p = sim.forward(c, a, 2, steps)
    w = rng.normal(size=p.shape)
    g = sim.backward(w)
    dt = sim.params.dt
    sim.close()
    px, py = [q[0] for q in PROBES], [q[1] for q in PROBES]

    def oracle_loss(aa):
        o = Oracle(nx=N_SIDE, ny=N_SIDE, dt=dt, **KW)
        o.set_fluid_state(*st)
        o.wavemaker_update(0.0, 0.0)
        o.time_step(1)
        rows = []
        for k in range(steps):
            it = 2 + k
            o.wavemaker_update(*paddle(aa, c.f0, dt, it))
            o.time_step(it)
            if (k + 1) % 5 == 0:
                rows.append(o.sampling_points(px, py, G)[:, 1])
        o.close()
        return float(np.sum(w * np.array(rows)))

    errs = []
    for j in range(3):
        d = rng.normal(size=8)
        d /= np.linalg.norm(d)
        fd, e, eps = fd_best(oracle_loss, a, d, g)
        say(f"oracle direction {j}: fd {fd:.10e} adjoint {g @ d:.10e} rel {e:.2e} (eps {eps})")
        errs.append(e)
    assert max(errs) < TOL, errs


def test_forward_rejects_what_the_reverse_does_not_cover(setup):
    st, a, _ = setup
    sim = make_sim(st, mls_on=1)
    with pytest.raises(sph2d.SphError, match="MLS"):
        sim.forward(cfg(), a, 2, 10)
    sim.close()
    sim = make_sim(st)
    with pytest.raises(sph2d.SphError, match="step 2"):
        sim.forward(cfg(), a, 1, 10)
    with pytest.raises(sph2d.SphError, match="completed"):
        sim.backward(np.zeros((2, 3)))
    sim.close()
