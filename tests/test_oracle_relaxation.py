"""Pins of the oracle's wave-theory wavemaker and its damping (relaxation) zone
(Wavemaker_Subroutine.f90:262-396,617-700; wave_maker_damping_zone.f90:61-95;
Parallel_Taylor_Green_Vortices_2D_Release.f90:113-143), each against what the
mathematics fixes rather than a retyped formula:

* zone membership is x <= L at set-up, the soft coefficient falls from 1 at the
  paddle to ~0 at L (exp(1.0) is REAL(4): -4.8e-8 there), monotonically;
* before the soft start ends (integer division makes the coefficient 0) the analytic
  fields are the hydrostatic rest state: zero velocity, initial position,
  p = rho0 g (d - y), rho from the inverse Tait law;
* the analytic velocity is the rate of change of the analytic position (linear
  regime, small wave height);
* Stokes_2nd differs from Airy by second-order terms: halving the wave height halves
  their relative difference;
* after the blend (shifting off) every zone particle satisfies rho = inverse Tait(p)
  to the last bit of pow and no particle outside does; the particle nearest the paddle (coefficient
  ~1) carries the analytic velocity;
* fluid below y = 0 is typed 120.
"""
import numpy as np

from cases import hydrostatic_tank
from oracle.pyoracle import Oracle

N, DEPTH, G = 40, 0.5, 9.81
OM = 4 * np.pi
PHASE = np.pi / 2


def wave_number(om=OM, d=DEPTH, g=G):
    k = om * om / g
    for _ in range(60):
        k = k - (g * k * np.tanh(k * d) - om * om) / (g * np.tanh(k * d) + g * k * d / np.cosh(k * d) ** 2)
    return k


def tank(zone=0.3, **kw):
    st = hydrostatic_tank(N, DEPTH, G)
    args = dict(integrator=2, walls=4, gravity=G, wavemaker_layers=3, water_depth=DEPTH, periodic=0,
                dt=1e-4 * 100 / N * 0.5)
    args.update(kw)
    o = Oracle(nx=N, ny=N, **args)
    o.set_fluid_state(*st)
    o.relaxation_init(zone)
    return o, st


def inv_tait(p, rho0=1000.0, c0sq=100.0):
    return rho0 * (1 + 7 * p / (c0sq * rho0)) ** float(np.float32(0.142857142))


def test_zone_membership_and_soft_coefficient():
    L = float(np.unique(hydrostatic_tank(N, DEPTH, G)[0])[11])   # a lattice column sits exactly at x = L
    o, st = tank(L)
    M = len(st[0])
    x = st[0]
    inz = o.i32("in_damping_zone")[:M]
    c = o.f64("position_soft_coefficient")[:M]
    assert np.array_equal(inz == 1, x <= L)
    xs = np.unique(x[x <= L])
    cs = np.array([c[x == v][0] for v in xs])
    assert np.all(np.diff(cs) < 0) and cs[0] > 0.99999 and cs[0] <= 1.0
    # at x = L exactly: exp((L/L)^3.5) = e, so 1 - (e - 1)/(e_REAL(4) - 1) = -4.80e-8 with the
    # correctly rounded REAL(4) e = 2.71828174591064453125
    e32 = float(np.float32(np.e))
    at_L = c[x == L]
    assert len(at_L) > 0 and np.all(at_L == 1 - (np.e - 1.0) / (e32 - 1.0)) and -1e-7 < at_L[0] < 0
    assert np.all(c[x > L] == 0.0)
    o.close()


def test_soft_start_gives_the_hydrostatic_rest_state():
    o, st = tank(0.3)
    M = len(st[0])
    k = wave_number()
    for wt in (1, 2):
        o.wave_theory_update(wt, 10, 0, o_dt(), G, DEPTH, 0.02, k, OM, PHASE, 3)   # 3 // 10 = 0: soft 0
        inz = o.i32("in_damping_zone")[:M] == 1
        x0, y0 = o.f64("initial_x")[:M], o.f64("initial_y")[:M]
        for f in ("an_u", "an_v"):
            assert np.all(o.f64(f)[:M][inz] == 0.0), f
        assert np.array_equal(o.f64("an_x")[:M][inz], x0[inz]) and np.array_equal(o.f64("an_y")[:M][inz], y0[inz])
        p = o.f64("an_p")[:M][inz]
        np.testing.assert_allclose(p, 1000.0 * G * (DEPTH - y0[inz]), rtol=1e-14, atol=1e-9)
        np.testing.assert_allclose(o.f64("an_rho")[:M][inz], inv_tait(p), rtol=1e-15)
        # the wavemaker particles rest at their initial positions, hydrostatic below the surface
        nwm = o.n_global - M - len(o.wall_positions()[0])
        gx, gy = o.f64("g_x"), o.f64("g_y")
        wm = slice(M, M + nwm)
        assert np.array_equal(gx[wm], o.f64("initial_x")[wm]) and np.all(o.f64("g_u")[wm] == 0.0)
        below = gy[wm] <= DEPTH
        np.testing.assert_allclose(o.f64("g_p")[wm][below], 1000.0 * G * (DEPTH - gy[wm][below]), rtol=1e-14,
                                   atol=1e-9)
        assert np.all(o.f64("g_mass")[wm][~below] == 0.0) and np.all(o.f64("g_p")[wm][~below] == 0.0)
    o.close()


_DT = 1e-4 * 100 / N * 0.5


def o_dt():
    return _DT


def test_analytic_velocity_is_the_rate_of_the_analytic_position():
    """Airy, soft coefficient 1, a small wave (k H/2 ~ 2e-3): the analytic position moves
    by the analytic velocity (central difference over two updates, after a few updates:
    the first one starts from the still-water position, and the depth factor of the
    next one jumps once by k zeta relative)."""
    o, st = tank(0.3)
    M = len(st[0])
    k = wave_number()
    H = 2e-4
    dt = 1e-3
    xs, us = [], []
    for it in range(10, 20):
        o.wave_theory_update(2, 1, 0, dt, G, DEPTH, H, k, OM, PHASE, it)
        xs.append(np.stack([o.f64("an_x")[:M], o.f64("an_y")[:M]]))
        us.append(np.stack([o.f64("an_u")[:M], o.f64("an_v")[:M]]))
    xs, us = xs[-3:], us[-3:]
    inz = o.i32("in_damping_zone")[:M] == 1
    rate = (xs[2] - xs[0]) / (2 * dt)
    vel = us[1]
    scale = np.max(np.abs(vel[:, inz]))
    assert scale > 1e-4
    assert np.max(np.abs(rate[:, inz] - vel[:, inz])) < 2e-2 * scale
    o.close()


def test_stokes_second_order_terms_scale_with_the_height():
    k = wave_number()
    diffs = []
    for H in (2e-3, 1e-3):
        out = {}
        for wt in (1, 2):
            o, st = tank(0.3)
            M = len(st[0])
            o.wave_theory_update(wt, 1, 0, 1e-3, G, DEPTH, H, k, OM, PHASE, 20)
            inz = o.i32("in_damping_zone")[:M] == 1
            out[wt] = o.f64("an_u")[:M][inz].copy()
            o.close()
        diffs.append(np.max(np.abs(out[1] - out[2])) / np.max(np.abs(out[2])))
    assert diffs[0] > 0
    ratio = diffs[1] / diffs[0]
    assert abs(ratio - 0.5) < 0.02, ratio


def test_blend_sets_density_from_pressure_in_the_zone_only():
    o, st = tank(0.3, shift_on=0, mls_on=0)
    M = len(st[0])
    k = wave_number()
    for it in range(1, 4):
        o.wave_theory_update(2, 1, 0, o_dt(), G, DEPTH, 0.02, k, OM, PHASE, it)
        o.time_step(it)
    inz = o.i32("in_damping_zone")[:M] == 1
    rho, p = o.f64("g_rho")[:M], o.f64("g_p")[:M]
    # numpy's power and C's pow differ by an ulp now and then
    np.testing.assert_allclose(rho[inz], inv_tait(p[inz]), rtol=5e-16, atol=0)
    off = np.abs(rho[~inz] - inv_tait(p[~inz])) / rho[~inz]
    assert np.count_nonzero(off < 1e-12) < 0.01 * np.count_nonzero(~inz)
    c = o.f64("position_soft_coefficient")[:M]
    j = np.argmax(np.where(inz, c, -1.0))
    u, au = o.f64("g_u")[:M], o.f64("an_u")[:M]
    assert abs(u[j] - au[j]) <= (1 - c[j]) * 10 * max(abs(au[j]), 1e-3)
    o.close()


def test_fluid_below_the_bed_is_typed_ill():
    o, st = tank(0.3)
    y = o.f64("g_y")
    y[5] = -1e-3
    o.set_f64("g_y", y)
    o.wave_theory_update(2, 1, 0, o_dt(), G, DEPTH, 0.02, wave_number(), OM, PHASE, 2)
    assert o.i32("g_type")[5] == 120 and o.i32("g_type")[6] == 2
    o.close()
