"""Reverse of one FP64 force evaluation (sph2d_single_step_compute_vjp).

The reverse has no counterpart in the reference; it is pinned by what fixes a
gradient: central finite differences of the forward evaluation, taken both on
the GPU forward itself and on the CPU oracle (an independent FP64 restatement
of single_step_compute.f90:62-171), along random directions in the state of
every fluid particle, and along single coordinates of particles whose periodic
ghosts take part (the ghost shares must come back to them). Tolerance 1e-6
relative to the directional derivative (or to a tenth of the sum of its terms'
sizes, where the terms cancel), 100x tighter than the 1e-4 that finite
differences of the whole step are asked to meet.
"""
import sys

import numpy as np
import pytest

import sph2d
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

N_SIDE = 40
TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noisy_state(n, seed):
    o = Oracle(nx=n, ny=n)
    o.init_tgv()
    N = n * n
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    x = o.f64("g_x")[:N] + rng.uniform(-0.25, 0.25, N) * dx
    y = o.f64("g_y")[:N] + rng.uniform(-0.25, 0.25, N) * dx
    u = o.f64("g_u")[:N] + rng.normal(0, 0.05, N)
    v = o.f64("g_v")[:N] + rng.normal(0, 0.05, N)
    rho = o.f64("g_rho")[:N] * (1 + rng.normal(0, 1e-3, N))
    m = o.f64("g_mass")[:N]
    return [x, y, u, v, rho], m


def equal_masses(m):
    """The SPS term reads particle_mass at the LOCAL index (Solver...SPS.f90:450-492),
    and the local order follows the positions, so with uneven masses a position
    perturbation can move the rates by a jump; equal masses make that read
    order-free, which lets finite differences run along positions."""
    return np.full_like(m, float(np.mean(m)))


def by_gid(name, vals, N):
    out = np.zeros(N)
    keep = name < N
    out[name[keep]] = vals[keep]
    return out


class GpuForward:
    def __init__(self, n, **kw):
        self.sim = sph2d.Simulation(sph2d.tgv_params(n, integrator=2, **kw))
        self.N = n * n

    def rates(self, state, m):
        s = self.sim
        s.upload_state(*state, m)
        s.distribute_period_boundary()
        s.devide_block(1)
        s.single_step_compute(1)
        name = s.local("name")
        out = [by_gid(name, s.local(f), self.N) for f in ("drhodt", "ax", "ay")]
        if s.params.epsilon_xsph != 0.0:
            out += [by_gid(name, s.local(f), self.N) for f in ("avg_x", "avg_y")]
        self.last_pairs = by_gid(name, s.local("neighbour_count").astype(np.float64), self.N)
        return out


class OracleForward:
    def __init__(self, n, **kw):
        self.n, self.kw, self.N = n, kw, n * n

    def rates(self, state, m):
        o = Oracle(nx=self.n, ny=self.n, integrator=2, dt=sph2d.tgv_params(self.n).dt, **self.kw)
        o.set_fluid_state(*state, m)
        o.period_boundary()
        o.devide_block(1)
        o.single_step_compute(1)
        name = o.i32("name")
        self.last_pairs = None
        return [by_gid(name, o.f64(f), self.N) for f in ("drhodt", "ax", "ay")]


def objective(rates, lam):
    return sum(float(np.dot(r, l)) for r, l in zip(rates, lam))


def directional_fd(fwd, state, m, lam, d, eps, base_pairs=None):
    """Central difference; None when the neighbour set changes between the two
    sides (a pair crossing the 3h support, where the kernel gradient jumps)."""
    plus = [s + eps * di for s, di in zip(state, d)]
    minus = [s - eps * di for s, di in zip(state, d)]
    jp = objective(fwd.rates(plus, m), lam)
    pp = fwd.last_pairs
    jm = objective(fwd.rates(minus, m), lam)
    pm = fwd.last_pairs
    if base_pairs is not None and not (np.array_equal(pp, base_pairs) and np.array_equal(pm, base_pairs)):
        return None
    return (jp - jm) / (2 * eps)


def scales(n):
    dx = 1.0 / n
    return [dx, dx, 0.05, 0.05, 1.0]     # x, y, u, v, rho (rho perturbed by 1 kg/m^3 per unit)


def reverse(sim, state, m, lam):
    sim.upload_state(*state, m)
    sim.distribute_period_boundary()
    sim.devide_block(1)
    sim.single_step_compute(1)
    g = sim.single_step_compute_vjp(*lam)
    return [g["x"], g["y"], g["u"], g["v"], g["rho"]]


@pytest.mark.parametrize("solver,xsph", [(sph2d.SOLVER_MODIFIED_LINEAR_RIEMANN, 0.0),
                                         (sph2d.SOLVER_LINEAR_RIEMANN, 0.0),
                                         (sph2d.SOLVER_MODIFIED_LINEAR_RIEMANN, 0.3)])
def test_reverse_evaluation_matches_finite_differences(solver, xsph):
    say(f"test_reverse_evaluation_matches_finite_differences solver={solver} xsph={xsph}")
    n, N = N_SIDE, N_SIDE * N_SIDE
    state, m_real = noisy_state(n, seed=11 + solver)
    m = equal_masses(m_real)
    rng = np.random.default_rng(3)
    nout = 5 if xsph else 3
    lam = [rng.normal(0, 1.0, N) for _ in range(nout)]
    gpu = GpuForward(n, solver=solver, epsilon_xsph=xsph)
    grad = reverse(gpu.sim, state, m, lam)
    again = reverse(gpu.sim, state, m, lam)
    for a, b in zip(grad, again):
        assert np.array_equal(a, b)      # fixed summation order: bit-identical
    sc = scales(n)
    gpu.rates(state, m)
    base = gpu.last_pairs
    orc = OracleForward(n, solver=solver, epsilon_xsph=xsph) if not xsph else None
    good = 0
    for k in range(10):
        d = [rng.normal(0, 1.0, N) * s for s in sc]
        adj = sum(float(np.dot(g, di)) for g, di in zip(grad, d))
        size = sum(float(np.sum(np.abs(g * di))) for g, di in zip(grad, d))   # before cancellation
        fd = directional_fd(gpu, state, m, lam, d, 1e-7, base)
        if fd is None:
            say(f"direction {k}: a pair crosses the support; skipped")
            continue
        say(f"direction {k}: reverse {adj:.12e} gpu fd {fd:.12e} rel {abs(adj - fd) / abs(fd):.2e}")
        assert abs(adj - fd) <= TOL * max(abs(fd), 0.1 * size), (k, adj, fd, size)
        if orc is not None and good < 2:
            fo = directional_fd(orc, state, m, lam[:3], d, 1e-7)
            say(f"   oracle fd {fo:.12e} rel {abs(adj - fo) / abs(fo):.2e}")
            assert abs(adj - fo) <= TOL * max(abs(fo), 0.1 * size), (k, adj, fo, size)
        good += 1
        if good == 4:
            break
    assert good == 4
    # the particles' own masses, along velocities and densities (no reorder there)
    grad = reverse(gpu.sim, state, m_real, lam)
    for k in range(2):
        d = [np.zeros(N), np.zeros(N)] + [rng.normal(0, 1.0, N) * s for s in sc[2:]]
        adj = sum(float(np.dot(g, di)) for g, di in zip(grad, d))
        size = sum(float(np.sum(np.abs(g * di))) for g, di in zip(grad, d))
        fd = directional_fd(gpu, state, m_real, lam, d, 1e-6)
        say(f"own masses, direction {k}: reverse {adj:.12e} gpu fd {fd:.12e} rel {abs(adj - fd) / abs(fd):.2e}")
        assert abs(adj - fd) <= TOL * max(abs(fd), 0.1 * size), (k, adj, fd, size)


def test_reverse_brings_ghost_shares_home():
    """Single coordinates of particles next to the periodic edges and in a corner:
    part of their gradient comes through their ghost copies."""
    say("test_reverse_brings_ghost_shares_home")
    n, N = N_SIDE, N_SIDE * N_SIDE
    state, m = noisy_state(n, seed=5)
    m = equal_masses(m)
    rng = np.random.default_rng(9)
    lam = [rng.normal(0, 1.0, N) for _ in range(3)]
    gpu = GpuForward(n)
    grad = reverse(gpu.sim, state, m, lam)
    gpu.rates(state, m)
    base = gpu.last_pairs
    x, y = state[0], state[1]
    corner = int(np.argmin(x + y))
    edge = int(np.argmin(np.abs(x - 0.5) + y))
    inner = int(np.argmin(np.abs(x - 0.5) + np.abs(y - 0.5)))
    sc = scales(n)
    for gid in (corner, edge, inner):
        for comp in range(5):
            d = [np.zeros(N) for _ in range(5)]
            d[comp][gid] = sc[comp]
            fd = directional_fd(gpu, state, m, lam, d, 1e-4, base)
            assert fd is not None, (gid, comp)
            adj = grad[comp][gid] * sc[comp]
            say(f"gid {gid} comp {comp}: reverse {adj:.10e} fd {fd:.10e}")
            # one coordinate moves the objective by little: rounding of the two
            # forward sums (~1e-9 of ~1e5) sets a 1e-5 bar here
            assert abs(adj - fd) <= 1e-5 * abs(fd), (gid, comp, adj, fd)


def test_reverse_rejects_what_it_does_not_cover():
    say("test_reverse_rejects_what_it_does_not_cover")
    n = 24
    state, m = noisy_state(n, seed=1)
    lam = [np.zeros(n * n)] * 3
    sim = sph2d.Simulation(sph2d.tgv_params(n, integrator=2, solver=sph2d.SOLVER_HLLC))
    with pytest.raises(sph2d.SphError):
        reverse(sim, state, m, lam)
    sim = sph2d.Simulation(sph2d.tgv_params(n, integrator=2))
    sim.upload_state(*state, m)
    with pytest.raises(sph2d.SphError):
        sim.single_step_compute_vjp(*lam)       # no evaluation yet


def test_vjp_refuses_a_state_changed_since_the_evaluation():
    """The reverse needs the local order, state and flags of the last evaluation:
    after a whole step, a fresh ordering or a shift it must refuse (SPH2D_ERR_STATE)
    instead of returning the gradient of another function."""
    say("test_vjp_refuses_a_state_changed_since_the_evaluation")
    n = 24
    st, m = noisy_state(n, 3)
    sim = sph2d.Simulation(sph2d.tgv_params(n, integrator=2))
    N = n * n
    lam = np.ones(N)

    def evaluate():
        sim.upload_state(*st, m)
        sim.distribute_period_boundary()
        sim.devide_block(1)
        sim.single_step_compute(1)

    evaluate()
    sim.single_step_compute_vjp(lam, lam, lam)            # accepted right after the evaluation
    for change in (lambda: sim.time_step(1), lambda: sim.devide_block(1),
                   lambda: sim.shift_particle_position(1), lambda: sim.refresh_domain(1)):
        evaluate()
        change()
        with pytest.raises(sph2d.SphError) as e:
            sim.single_step_compute_vjp(lam, lam, lam)
        assert e.value.code == sph2d.ERR_STATE
    evaluate()
    with pytest.raises(ValueError):
        sim.single_step_compute_vjp(lam[:-1], lam, lam)   # a weight array of the wrong length
    sim.close()


def test_device_array_reverse_equals_the_host_array_one():
    """sph2d_single_step_compute_vjp_device packs the weights and unpacks the
    gradients on the device: with the same weights it returns the host-array
    entry's gradients bit for bit, with and without XSPH weights."""
    say("test_device_array_reverse_equals_the_host_array_one")
    n = 32
    st, m = noisy_state(n, 11)
    sim = sph2d.Simulation(sph2d.tgv_params(n, integrator=2))
    N = n * n
    rng = np.random.default_rng(13)
    lam = [rng.normal(0, 1.0, N) for _ in range(5)]
    sim.upload_state(*st, m)
    sim.distribute_period_boundary()
    sim.devide_block(1)
    sim.single_step_compute(1)
    for with_avg in (False, True):
        host = sim.single_step_compute_vjp(*lam) if with_avg else sim.single_step_compute_vjp(*lam[:3])
        d_ld = sph2d.DeviceArray(lam[0])
        d_la = sph2d.DeviceArray(np.concatenate([lam[1], lam[2]]))
        d_lg = sph2d.DeviceArray(np.concatenate([lam[3], lam[4]])) if with_avg else None
        gp, gv, gr = (sph2d.DeviceArray(np.full(k * N, np.nan)) for k in (2, 2, 1))
        sim.single_step_compute_vjp_device(d_ld, d_la, gp, gv, gr, lam_avg=d_lg)
        p, v, r = gp.get(), gv.get(), gr.get()
        for k, dev in (("x", p[:N]), ("y", p[N:]), ("u", v[:N]), ("v", v[N:]), ("rho", r)):
            assert np.array_equal(dev, host[k]), (with_avg, k)
        say(f"xsph weights {with_avg}: identical")
    with pytest.raises(ValueError):
        sim.single_step_compute_vjp_device(sph2d.DeviceArray(lam[0][:-1]), d_la, gp, gv, gr)
    sim.close()


def test_reverse_matches_finite_differences_around_a_free_surface():
    """A disc of particles removed from the noisy periodic lattice: particles on the
    hole's rim are free-surface particles and their neighbours near-surface ones, so
    the near-surface branch of the pair terms (v* = v_bar, P* = P_bar) and the
    surface-gated terms enter the reverse; checked against central differences of
    the GPU forward and of the oracle, with the flags held as the forward sets them
    (directions that change them, or the neighbour set, are skipped)."""
    say("test_reverse_matches_finite_differences_around_a_free_surface")
    n = N_SIDE
    state, m_real = noisy_state(n, seed=21)
    keep = (state[0] - 0.5) ** 2 + (state[1] - 0.5) ** 2 > 0.2 ** 2
    state = [a[keep] for a in state]
    m = equal_masses(m_real[keep])
    N = int(keep.sum())
    rng = np.random.default_rng(17)
    lam = [rng.normal(0, 1.0, N) for _ in range(3)]
    gpu = GpuForward(n)
    gpu.N = N
    gpu.rates(state, m)
    surf = gpu.sim.local("surface")
    near = gpu.sim.local("near_surface")
    say(f"{N} particles, {int(surf.sum())} surface, {int(near.sum())} near-surface slots")
    assert surf.sum() > 0 and near.sum() > 0
    base = gpu.last_pairs
    grad = reverse(gpu.sim, state, m, lam)
    sc = scales(n)
    orc = OracleForward(n)
    orc.N = N
    good = 0
    for k in range(12):
        d = [rng.normal(0, 1.0, N) * s for s in sc]
        adj = sum(float(np.dot(g, di)) for g, di in zip(grad, d))
        size = sum(float(np.sum(np.abs(g * di))) for g, di in zip(grad, d))
        fd = directional_fd(gpu, state, m, lam, d, 1e-7, base)
        if fd is None or not np.array_equal(gpu.sim.local("surface"), surf):
            say(f"direction {k}: the neighbour set or the flags change; skipped")
            continue
        say(f"direction {k}: reverse {adj:.12e} gpu fd {fd:.12e} rel {abs(adj - fd) / abs(fd):.2e}")
        assert abs(adj - fd) <= TOL * max(abs(fd), 0.1 * size), (k, adj, fd, size)
        if good == 0:
            fo = directional_fd(orc, state, m, lam, d, 1e-7)
            say(f"   oracle fd {fo:.12e} rel {abs(adj - fo) / abs(fo):.2e}")
            assert abs(adj - fo) <= TOL * max(abs(fo), 0.1 * size), (k, adj, fo, size)
        good += 1
        if good == 3:
            break
    assert good == 3
    gpu.sim.close()
