"""Pins the oracle's HLLC and MUSCL solver variants
(Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90,
Solver_Normalized_MUSCL_Linear_Riemann_Delta_SPS.f90).

* MUSCL_Limiter is compared bit for bit with the reference's own routine
  (function_module.f90:34-80, compiled into oracle/_ref).
* The HLLC star state is checked against what the mathematics fixes: the
  acoustic star state solves the two linear wave relations
  P* - P_L = -z_L (U* - U_L), P* - P_R = z_R (U* - U_R) (solved here with numpy).
  U* is that solution; the reference's P* (:383) weights each side's pressure by
  its OWN impedance, z_L P_L + z_R P_R, where the acoustic solution has
  z_R P_L + z_L P_R: the two agree for z_L = z_R and differ by exactly
  (z_L - z_R)(P_L - P_R)/(z_L + z_R) otherwise, which the test states;
  the tangential velocity is the left one left of the contact and the right one
  right of it; far outside the fan the sample is the undisturbed state; S = U -/+ c
  where both waves are rarefactions (P* below both pressures).
* The shock-side wave-speed factor Q (:388-412) is pinned only through its
  weak-shock limit Q^2 -> 7 P_L / (rho_L c_L^2): parity partially pinned.
* The MUSCL gradients (:145-185) are checked against analytic gradients of smooth
  periodic fields.
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

D = ctypes.c_double


def ref_limiter(di, dj, dij):
    ref = po.ref_lib()
    li, lj = D(), D()
    ref._QMfunction_modulePmuscl_limiter(ctypes.byref(D(di)), ctypes.byref(D(dj)), ctypes.byref(D(dij)),
                                         ctypes.byref(li), ctypes.byref(lj))
    return li.value, lj.value


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_muscl_limiter_bit_exact_vs_reference():
    rng = np.random.default_rng(3)
    cases = [(0.0, 0.0, 0.0), (1.0, 1.0, 2.0), (1.0, 1.0, 1.0), (1.0, -1.0, 2.0), (-1.0, 1.0, -2.0),
             (0.5, 0.7, 0.1), (-0.5, -0.7, -0.1), (0.7, 0.5, 0.1), (-0.7, -0.5, -0.1), (2.0, 1.0, 5.0)]
    for _ in range(20000):
        a, b = rng.normal(size=2) * 10.0 ** rng.uniform(-3, 3, 2)
        d = rng.normal() * 10.0 ** rng.uniform(-3, 3)
        if rng.uniform() < 0.2:
            b = a * rng.choice([1.0, 1.3, 1 / 1.3, -1.0])
        if rng.uniform() < 0.1:
            d = 2 * a
        cases.append((float(a), float(b), float(d)))
    n_branch = {"pass": 0, "clip": 0}
    for di, dj, dij in cases:
        r = ref_limiter(di, dj, dij)
        m = po.muscl_limiter(di, dj, dij)
        assert r == m, (di, dj, dij, r, m)
        n_branch["clip" if m[0] != m[1] else "pass"] += 1
    assert min(n_branch.values()) > 1000


def test_muscl_limiter_zero_at_opposite_slopes_and_bounded():
    rng = np.random.default_rng(4)
    for _ in range(2000):
        a, b, d = rng.normal(size=3)
        li, lj = po.muscl_limiter(a, -abs(b) * np.sign(a), d)   # slopes of opposite sign
        assert li == 0.0 and lj == 0.0
        li, lj = po.muscl_limiter(a, b, d)
        assert abs(li) <= 1.3 * max(abs(a), abs(b)) + abs(d) and abs(lj) <= 1.3 * max(abs(a), abs(b)) + abs(d)


def _rot(e, v):
    """(normal, tangential) components of v in the frame of unit vector e."""
    return np.array([v[0] * e[0] + v[1] * e[1], -v[0] * e[1] + v[1] * e[0]])


def _acoustic(UL, UR, PL, PR, zL, zR):
    A = np.array([[zL, 1.0], [-zR, 1.0]])
    b = np.array([zL * UL + PL, -zR * UR + PR])
    return np.linalg.solve(A, b)                  # (U*, P*)


@pytest.mark.parametrize("seed", range(6))
def test_hllc_star_states_are_the_acoustic_solution(seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi)
    e = np.array([np.cos(th), np.sin(th)])
    rhoL, rhoR = 1000 * (1 + rng.normal(0, 1e-3, 2))
    cL, cR = 10 * (1 + rng.normal(0, 1e-2, 2))
    PL, PR = 200 + rng.normal(0, 20, 2)
    vl = rng.normal(0, 1, 2)
    vr = vl.copy()
    # make it an expansion (U_L < U_R): P* below both pressures, Q = 1 on both sides
    vr += e * (2.0 + abs(PL - PR) / (rhoL * cL))
    L, R = _rot(e, vl), _rot(e, vr)
    Us, Ps = _acoustic(L[0], R[0], PL, PR, rhoL * cL, rhoR * cR)
    assert Ps < min(PL, PR)
    SL, SR = L[0] - cL, R[0] + cR
    # left star region: S_L <= x/t <= S*
    vs, ps = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, 0.5 * (SL + Us))
    zL, zR = rhoL * cL, rhoR * cR
    np.testing.assert_allclose(ps, Ps + (zL - zR) * (PL - PR) / (zL + zR), rtol=1e-12)
    _, ps_eq = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoL, cL, e, 0.5 * (SL + Us))
    np.testing.assert_allclose(ps_eq, _acoustic(L[0], R[0], PL, PR, zL, zL)[1], rtol=1e-12)
    np.testing.assert_allclose(_rot(e, vs), [Us, L[1]], rtol=1e-12, atol=1e-12)
    # right star region: S* <= x/t <= S_R
    vs, ps = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, 0.5 * (SR + Us))
    np.testing.assert_allclose(_rot(e, vs), [Us, R[1]], rtol=1e-12, atol=1e-12)
    # wave speeds: just outside S_L the left state, just inside the left star state
    vs, _ = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, SL - 1e-9)
    np.testing.assert_allclose(vs, vl, rtol=1e-13, atol=1e-14)
    vs, _ = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, SL + 1e-9)
    np.testing.assert_allclose(_rot(e, vs)[0], Us, rtol=1e-12)
    vs, _ = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, SR + 1e-9)
    np.testing.assert_allclose(vs, vr, rtol=1e-13, atol=1e-14)
    vs, _ = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, SR - 1e-9)
    np.testing.assert_allclose(_rot(e, vs)[0], Us, rtol=1e-12)
    # x/t = dx/dt of the lattice runs (100) lies right of the fan: the right state
    vs, _ = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, 100.0)
    np.testing.assert_allclose(vs, vr, rtol=1e-13, atol=1e-14)
    # the star pressure does not depend on which particle is called left
    _, ps2 = po.hllc_star(vr, PR, rhoR, cR, vl, PL, rhoL, cL, -e, 0.0)
    np.testing.assert_allclose(ps2, ps, rtol=1e-13)


def test_hllc_weak_shock_wave_speed_limit():
    """Compression (U_L > U_R) with P* just above P_L: the left wave speed found by
    bisection on x/t tends to U_L - c_L sqrt(7 P_L / (rho_L c_L^2))."""
    e = np.array([1.0, 0.0])
    rhoL = rhoR = 1000.0
    cL = cR = 10.0
    PL = PR = 300.0
    for dU in (1e-4, 1e-5):
        vl = np.array([dU, 0.3])
        vr = np.array([0.0, -0.2])
        L = _rot(e, vl)
        Us, Ps = _acoustic(L[0], 0.0, PL, PR, rhoL * cL, rhoR * cR)
        assert Ps > PL
        lo, hi = -1e3, Us - 1e-12            # left state at lo, left star at hi
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            vs, _ = po.hllc_star(vl, PL, rhoL, cL, vr, PR, rhoR, cR, e, mid)
            if vs[1] == vl[1] and vs[0] == vl[0]:
                lo = mid
            else:
                hi = mid
        SL_expected = L[0] - cL * np.sqrt(7 * PL / (rhoL * cL * cL))
        assert abs(lo - SL_expected) < 1e-3 * abs(SL_expected), (lo, SL_expected)


def _evaluate(o, it=1):
    o.period_boundary()
    o.devide_block(it)
    o.single_step_compute(it)
    return o


def test_muscl_gradients_match_analytic_fields():
    n = 64
    o = po.Oracle(nx=n, ny=n, solver=po.ORC_SOLVER_MUSCL)
    o.init_tgv()
    N = n * n
    x, y = o.f64("g_x")[:N].copy(), o.f64("g_y")[:N].copy()
    k = 2 * np.pi
    u = 0.5 * np.sin(k * x) * np.cos(k * y)
    v = 0.3 * np.cos(2 * k * x)
    rho = 1000.0 * (1 + 1e-3 * np.sin(k * y))
    m = rho / N
    o.set_fluid_state(x, y, u, v, rho, m)
    _evaluate(o)
    t = o.i32("type") == 2
    xs, ys = o.f64("x")[t], o.f64("y")[t]
    us_x = 0.5 * k * np.cos(k * xs) * np.cos(k * ys)
    us_y = -0.5 * k * np.sin(k * xs) * np.sin(k * ys)
    vs_x = -0.3 * 2 * k * np.sin(2 * k * xs)
    rs = 1000.0 * (1 + 1e-3 * np.sin(k * ys))
    B = 100.0 * 1000.0 / 7
    ps_y = 7 * B / 1000.0 * (rs / 1000.0) ** 6 * (1000.0 * 1e-3 * k * np.cos(k * ys))
    for got, want in ((o.f64("grad_u_x")[t], us_x), (o.f64("grad_u_y")[t], us_y),
                      (o.f64("grad_v_x")[t], vs_x), (o.f64("grad_p_y")[t], ps_y)):
        assert np.max(np.abs(got - want)) < 0.03 * np.max(np.abs(want))
    assert np.max(np.abs(o.f64("grad_v_y")[t])) < 1e-9 * np.max(np.abs(vs_x))
    assert np.max(np.abs(o.f64("grad_p_x")[t])) < 1e-9 * np.max(np.abs(ps_y))
