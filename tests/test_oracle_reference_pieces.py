"""Pins the C oracle bit-for-bit against the reference's own Fortran routines,
compiled from the reference sources by oracle/build_ref.sh (flang, no FMA):
compute_kernel.f90, status_equation.f90, function_module.f90 (BRINV,
matrix_inversion_less_three) and information_module.f90's constants.
Skipped where oracle/_ref is not built (where the reference sources are absent)."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (reference sources absent)")


class _Lazy:
    """The compiled reference, loaded on first use (never at collection)."""

    def __getattr__(self, name):
        return getattr(po.ref_lib(), name)


ref = _Lazy()

D = ctypes.c_double
I = ctypes.c_int32


def ref_kernel(r, pd, h):
    w = D()
    dw = (D * 2)()
    pdx = (D * 2)(*pd)
    ref.compute_kernel_(ctypes.byref(I(2)), ctypes.byref(D(r)), pdx, ctypes.byref(D(h)), ctypes.byref(w), dw)
    return w.value, np.array([dw[0], dw[1]])


def test_kernel_bit_exact():
    rng = np.random.default_rng(7)
    for h in (0.012, 1.2 / 256, 1.2 / 2048, 0.0115):
        for _ in range(400):
            r = rng.uniform(1e-7, 3.2 * h)
            th = rng.uniform(0, 2 * np.pi)
            pd = np.array([r * np.cos(th), r * np.sin(th)])
            r = float(np.sqrt(pd[0] * pd[0] + pd[1] * pd[1]))
            w0, d0 = ref_kernel(r, pd, h)
            w1, d1 = po.compute_kernel(r, pd, h)
            assert w0 == w1 and np.array_equal(d0, d1), (h, r, w0, w1, d0, d1)


def test_status_equation_bit_exact():
    o = po.Oracle()
    L = po.lib()
    rng = np.random.default_rng(3)
    for rho in 1000.0 * (1 + rng.uniform(-0.02, 0.02, 2000)):
        p, c = D(), D()
        ref.status_equation_(ctypes.byref(D(rho)), ctypes.byref(D(0.0)), ctypes.byref(I(2)),
                             ctypes.byref(p), ctypes.byref(c))
        p1, c1 = D(), D()
        L.orc_status_equation(o._c, rho, 0.0, 2, ctypes.byref(p1), ctypes.byref(c1))
        assert p.value == p1.value and c.value == c1.value


def _ref_brinv(a):
    m = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    l = I(0)
    ref._QMfunction_modulePbrinv(m.ctypes.data_as(ctypes.POINTER(D)), ctypes.byref(I(a.shape[0])), ctypes.byref(l))
    return l.value, m


def test_brinv_bit_exact_and_correct():
    rng = np.random.default_rng(11)
    for n in (2, 3, 6):
        for _ in range(200):
            a = rng.normal(size=(n, n))
            if n == 6:  # MLS-shaped: sum of w p p^T
                p = rng.normal(size=(6, 30))
                a = (p * rng.uniform(0, 1, 30)) @ p.T
            l0, m0 = _ref_brinv(a)
            l1, m1 = po.brinv(a)
            assert l0 == l1 == 1
            assert np.array_equal(m0, m1)
            np.testing.assert_allclose(m1 @ a, np.eye(n), atol=1e-9 * np.abs(m1).max() * np.abs(a).max())
    z = np.zeros((6, 6))
    assert _ref_brinv(z)[0] == 0 and po.brinv(z)[0] == 0


def test_matrix_inversion_less_three():
    rng = np.random.default_rng(5)
    f = ref._QMfunction_modulePmatrix_inversion_less_three
    for n in (2, 3):
        for _ in range(200):
            a = np.asfortranarray(rng.normal(size=(n, n)))
            res = np.zeros((n, n), order="F")
            chk = I(0)
            f(ctypes.byref(I(n)), a.ctypes.data_as(ctypes.POINTER(D)), res.ctypes.data_as(ctypes.POINTER(D)),
              ctypes.byref(chk))
            ok, mine = po.matrix_inversion_less_three(a)
            assert chk.value == ok == 1
            assert np.array_equal(res, mine)
            np.testing.assert_allclose(mine, np.linalg.inv(a), rtol=1e-10, atol=1e-12)
    # singular: check_error = 0 and only the diagonal is preset to 1
    ok, mine = po.matrix_inversion_less_three(np.ones((2, 2)))
    assert ok == 0 and mine[0, 0] == 1.0 and mine[1, 1] == 1.0


def test_constants_match_information_module():
    o = po.Oracle()
    assert po.ref_constant("smooth_length") == 1.2 * (1.0 / 100)
    assert po.ref_constant("water_kinematic_viscosity") == o.params.nu
    assert po.ref_constant("square_c_0") == o.params.c0sq
    assert po.ref_constant("chain_dx") == 4 * 1.2 * (1.0 / 100)
    assert po.ref_constant("chain_origin_x") == -(5 * (4 * 1.2 * (1.0 / 100)))
    assert po.ref_int_constant("chain_x_number") == int(1.0 / (4 * 1.2 * 0.01) + 10) == 30
    assert po.ref_constant("max_shift") == 0.25 * 0.01
    assert po.ref_constant("water_rho_0") == o.params.rho0
    assert po.ref_constant("reynolds_number") == 1.0 * 1.0 / o.params.nu
