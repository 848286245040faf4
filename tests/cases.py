"""Initial states shared by the oracle tests and the GPU parity tests."""
import numpy as np

from oracle import pyoracle as po


def hydrostatic_tank(n=30, H=0.5, g=9.81, rho0=1000.0, c0sq=100.0):
    """Fluid at rest below y = H in the walled unit box, p = rho0 g (H - y), rho from
    the inverse Tait law (REAL(4) exponent, Distribute_Initial_Particle.f90:222)."""
    o0 = po.Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    x, y = o0.f64("g_x")[:N].copy(), o0.f64("g_y")[:N].copy()
    o0.close()
    keep = y < H
    x, y = x[keep], y[keep]
    p = rho0 * g * (H - y)
    rho = rho0 * (1 + 7 * p / (c0sq * rho0)) ** float(np.float32(0.142857142))
    M = len(x)
    return x, y, np.zeros(M), np.zeros(M), rho, rho / n ** 2
