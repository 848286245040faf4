"""Pins of the oracle's 2x2 renormalisation mode (renorm = 1, a mode of this library:
the reference forms T_i = sum_j (x_j - x_i) (x) grad_i W V_j and its inverse,
Define_freesurface_and_Normailized_derivates.f90:141-188, then uses the identity):

* consistency, the point of the correction: for a linear density field the
  normalised grad rho of Solver...SPS.f90:508-526 equals the field's gradient at
  every particle, surface particles of a holed disc included, where the identity's
  truncated support misses it by tens of percent;
* the tensor is symmetric, and diagonal and close to the identity inside a full
  lattice support;
* matrix_inversion_less_three's singular case (det = 0 exactly) keeps the identity
  and is counted: a particle whose neighbours all lie on one horizontal line has
  T_yy = 0 = T_xy, and an isolated particle has T = 0.
"""
import numpy as np

import oracle.pyoracle as po
from test_oracle_pins import holed_disc, run_eval


def linear_field_state(n=60, a=(30.0, -20.0)):
    x, y, u, v, rho, m = holed_disc(n)
    rho = 1000.0 * (1.0 + a[0] * 1e-3 * x + a[1] * 1e-3 * y)
    return (x, y, u, v, rho, m), np.array(a) * 1e-3 * 1000.0


def grad_rho(renorm, n=60):
    st, g = linear_field_state(n)
    o = po.Oracle(nx=n, ny=n, integrator=2, renorm=renorm)
    o.set_fluid_state(*st)
    run_eval(o)
    name = o.i32("name")
    fluid = name < len(st[0])
    gr = np.stack([o.f64("grad_rho_x"), o.f64("grad_rho_y")], axis=1)[fluid]
    surf = o.i32("surface")[fluid].astype(bool)
    T = o.f64("T").reshape(-1, 4)[fluid]
    o.close()
    return gr, g, surf, T


def test_renormalised_gradient_reproduces_a_linear_field():
    gr1, g, surf, T = grad_rho(1)
    err1 = np.max(np.linalg.norm(gr1 - g, axis=1)) / np.linalg.norm(g)
    assert err1 < 1e-12, err1
    gr0, _, _, _ = grad_rho(0)
    err0 = np.linalg.norm(gr0 - g, axis=1) / np.linalg.norm(g)
    assert surf.sum() > 20 and np.max(err0[surf]) > 0.2


def test_tensor_is_symmetric_and_near_identity_inside():
    _, _, surf, T = grad_rho(1)
    # column-major T(L, o): entries 1 and 2 are T(1,0) and T(0,1)
    np.testing.assert_allclose(T[:, 1], T[:, 2], rtol=0, atol=1e-13)
    # a full support on the lattice (periodic Taylor-Green box): T = I to the quadrature
    # error of the truncated kernel, diagonal, and the same for every particle
    o = po.Oracle(nx=40, ny=40, integrator=2, renorm=1)
    o.init_tgv()
    run_eval(o)
    name = o.i32("name")
    Tl = o.f64("T").reshape(-1, 4)[name < 1600]
    o.close()
    assert np.max(np.abs(Tl[:, [0, 3]] - 1.0)) < 5e-3 and np.max(np.abs(Tl[:, [1, 2]])) < 1e-12
    assert np.ptp(Tl[:, 0]) < 1e-12


def test_singular_tensors_keep_the_identity_and_are_counted():
    n = 40
    dx = 1.0 / n
    pts = []
    # 6 horizontal pairs, each 1 dx long and far from the rest: T_yy = T_xy = 0
    for k in range(6):
        x0, y0 = 0.1 + 0.13 * k, 0.2
        pts += [(x0, y0), (x0 + dx, y0)]
    # 3 isolated particles: T = 0
    pts += [(0.2, 0.6), (0.5, 0.6), (0.8, 0.6)]
    # a 5x5 block: full-rank tensors
    for i in range(5):
        for j in range(5):
            pts.append((0.4 + i * dx, 0.85 + j * dx * 0.5))
    p = np.array(pts)
    k = len(p)
    rho = np.full(k, 1000.0)
    o = po.Oracle(nx=n, ny=n, integrator=2, renorm=1)
    o.set_fluid_state(p[:, 0], p[:, 1], np.zeros(k), np.zeros(k), rho, rho * dx * dx)
    run_eval(o)
    assert o.n_singular == 15
    name = o.i32("name")
    Ti = o.f64("T_inv").reshape(-1, 4)
    single = name < 15
    assert np.all(Ti[single] == np.array([1.0, 0.0, 0.0, 1.0]))
    assert not np.all(Ti[~single] == np.array([1.0, 0.0, 0.0, 1.0]))
    o.close()
