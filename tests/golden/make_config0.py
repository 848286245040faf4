"""Writes tests/golden/tgv100_pc_rho1.npz: the oracle's Taylor-Green run as the
reference's runnable configuration states it -- 100^2 lattice, rho0 = 1, U = 1,
Re = 100, c0 = 10, predictor-corrector at the CFL time step of
Initialize_SPH_Calculation.f90:449-461 (dt = 5.4545e-5), to t = 0.5 (9,167 steps).
The run takes minutes on the host, too long for the test suite, so its samples
(Sampling_subroutine.f90:340-461: KE/KE0, max|v|, the error norms) are kept every
1,000 steps and at the last step.

    python tests/golden/make_config0.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle.pyoracle import Oracle  # noqa: E402

N = 100
RECORD_EVERY = 1000


def cfl_dt(h, c0):
    """2 x 0.1 x min(Morris 0.125 h^2 / 1.01e-6, Monaghan 0.25 h / (1.1 c0)), the
    literals REAL(4) as in the Fortran."""
    f32 = lambda v: float(np.float32(v))
    morris = 0.125 * h * h / f32(1.01e-6)
    monaghan = 0.25 * h / (f32(1.1) * math.sqrt(c0 * c0))
    return float(np.float32(2) * np.float32(0.1)) * min(morris, monaghan)


def main():
    dt = cfl_dt(1.2 / N, 10.0)
    steps = int(round(0.5 / dt))
    o = Oracle(nx=N, ny=N, integrator=2, dt=dt, rho0=1.0)
    o.init_tgv()
    rows = []
    for it in range(1, steps + 1):
        o.time_step(it)
        if it % RECORD_EVERY == 0 or it == steps:
            rows.append([it, it * dt] + o.sample().tolist())
            print(rows[-1][:3], flush=True)
    o.close()
    np.savez(os.path.join(HERE, "tgv100_pc_rho1.npz"), dt=dt, samples=np.array(rows))


if __name__ == "__main__":
    main()
