"""Writes tests/golden/tgv_reference.npz from the reference's own shipped output
files (run in the development container, where /root/reference exists).

Sources (mini-walker/SPH-2D, Taylor_Green_Vortex/, a 5-rank MPI run of the
default case: 100x100 TGV, RK-variant, dt=1e-4):
  Kinetic_energy.dat              t, KE/KE0                 (500 rows, every 100 steps)
  Maximum_velocity_magnitude.dat  t, max|v|/U
  Norm_Horizontal_velocity.dat    t, Linf, L1, L2 of (u-u_a)/U
  Norm_Vertical_velocity.dat      t, Linf, L1, L2 of (v-v_a)/U
  Norm_Pressure.dat               t, Linf, L1, L2 of (p-p_a)/(rho0 U^2)
  Norm_Velocity_magnitude.dat     t, Linf, L1, L2 of (|v|-|v_a|)/U
  Initial_particle_information.dat  x y u v rho p/(rho0 U^2) at t=0 (F20.10), every 10th row kept
  Free_Surface.dat                zone sizes of the free-surface dumps
Only data is copied; no reference source is.
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("SPH2D_REFERENCE", "/root/reference") + "/Taylor_Green_Vortex"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tgv_reference.npz")


def table(name):
    return np.loadtxt(os.path.join(REF, name))


def particles(name):
    rows = []
    with open(os.path.join(REF, name)) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 6:
                try:
                    rows.append([float(v) for v in parts])
                except ValueError:
                    pass
    return np.array(rows)


def main():
    if not os.path.isdir(REF):
        print("reference not present; nothing to do")
        return 0
    init = particles("Initial_particle_information.dat")
    assert init.shape == (10000, 6), init.shape
    keep = np.arange(0, 10000, 10)
    fs = []
    with open(os.path.join(REF, "Free_Surface.dat")) as f:
        for line in f:
            m = re.search(r"ZONE.*\bI=\s*(\d+)", line)
            if m:
                fs.append(int(m.group(1)))
    np.savez_compressed(
        OUT,
        ke=table("Kinetic_energy.dat"),
        vmax=table("Maximum_velocity_magnitude.dat"),
        norm_u=table("Norm_Horizontal_velocity.dat"),
        norm_v=table("Norm_Vertical_velocity.dat"),
        norm_p=table("Norm_Pressure.dat"),
        norm_vm=table("Norm_Velocity_magnitude.dat"),
        init_index=keep,
        init=init[keep],
        free_surface_zone_sizes=np.array(fs, dtype=np.int64),
    )
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    return 0


if __name__ == "__main__":
    sys.exit(main())
