"""The wave-theory wavemaker and its damping (relaxation) zone on the GPU against the
oracle (Wavemaker_Subroutine.f90:262-396,617-700; wave_maker_damping_zone.f90:61-95;
time_iteration.f90:660-667,1052-1059): a walled tank of still water 0.5 deep (40^2
lattice), three columns of wavemaker particles moved by Airy or Stokes_2nd theory,
the zone over x <= 0.3, gravity; 12 steps of either integrator, the state after
them within the parity bars (positions 1e-12 absolute, the rest 1e-10 of the field).

The wave kinematics are evaluated on each side (device sin/cos/sinh/cosh and glibc's
round an ulp apart now and then): the states agree to ~1e-16 in position for the
first steps; the wavemaker's start (the soft-start coefficient jumps from 0 to 1,
Wavemaker_Subroutine.f90:283-287 divides integers) drives the free surface hard
enough that, some 15-20 steps in, an ulp flips a surface flag and the two runs part
at ~1e-7 (with or without the zone). The comparison stops before that."""
import numpy as np
import pytest

import sph2d
from cases import hydrostatic_tank
from oracle.pyoracle import Oracle
from test_gpu_parity import rel_err, say

pytestmark = pytest.mark.gpu

N, DEPTH, G = 40, 0.5, 9.81
OM = 4 * np.pi


def wave_number(om=OM, d=DEPTH, g=G):
    k = om * om / g
    for _ in range(60):
        k = k - (g * k * np.tanh(k * d) - om * om) / (g * np.tanh(k * d) + g * k * d / np.cosh(k * d) ** 2)
    return k


def run_pair(wave_type, integrator, zone, steps, soft_steps=1, **extra):
    st = hydrostatic_tank(N, DEPTH, G)
    M = len(st[0])
    kw = dict(integrator=integrator, walls=4, gravity=G, wavemaker_layers=3, water_depth=DEPTH)
    kw.update(extra)
    prm = sph2d.tgv_params(N, periodic=0, dt=1e-4 * 100 / N * 0.5, **kw)
    sim = sph2d.Simulation(prm, capacity=N * N)
    sim.upload_state(*st)
    o = Oracle(nx=N, ny=N, dt=prm.dt, periodic=0, **kw)
    o.set_fluid_state(*st)
    sim.relaxation_init(zone)
    o.relaxation_init(zone)
    k = wave_number()
    w = sph2d.Wavemaker(wave_type=wave_type, soft_steps=soft_steps, standing_steps=0, dt=prm.dt, g=G, water_depth=DEPTH,
                        wave_height=0.02, wave_number=k, wave_frequency=OM, phase=np.pi / 2)
    for it in range(1, steps + 1):
        sim.wave_theory_update(w, it)
        o.wave_theory_update(wave_type, soft_steps, 0, prm.dt, G, DEPTH, 0.02, k, OM, np.pi / 2, it)
        sim.time_step(it)
        o.time_step(it)
    return sim, o, M


@pytest.mark.parametrize("wave_type,integrator", [(sph2d.WAVE_AIRY, 2), (sph2d.WAVE_STOKES2, 2),
                                                  (sph2d.WAVE_AIRY, 3)])
def test_wave_theory_tank_with_damping_zone_parity(wave_type, integrator):
    say(f"test_wave_theory_tank_with_damping_zone_parity wave_type={wave_type} integrator={integrator}")
    sim, o, M = run_pair(wave_type, integrator, 0.3, 12)
    g = sim.download_state()
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v"), ("rho", "g_rho"), ("p", "g_p")):
        r = o.f64(of)[:M]
        e = float(np.max(np.abs(g[f] - r))) if f in ("x", "y") else rel_err(g[f], r)
        say(f, e)
        assert e < (1e-12 if f in ("x", "y") else 1e-10), (f, e)
    # the zone acted: zone particles differ from a run without it by far more than the bar
    free, fo, _ = run_pair(wave_type, integrator, 0.0, 12)
    gf = free.download_state()
    inz = o.i32("in_damping_zone")[:M] == 1
    assert np.max(np.abs(gf["u"][inz] - g["u"][inz])) > 1e-6
    assert np.array_equal(o.i32("g_type")[:M], g["type"])
    for s in (sim, free):
        s.close()
    o.close()
    fo.close()


def test_wave_theory_entries_check_their_state():
    st = hydrostatic_tank(N, DEPTH, G)
    prm = sph2d.tgv_params(N, periodic=0, walls=4, gravity=G, wavemaker_layers=3, water_depth=DEPTH, integrator=2)
    sim = sph2d.Simulation(prm, capacity=N * N)
    sim.upload_state(*st)
    w = sph2d.Wavemaker(wave_type=sph2d.WAVE_AIRY, soft_steps=1, standing_steps=0, dt=prm.dt, g=G,
                        water_depth=DEPTH, wave_height=0.02, wave_number=wave_number(), wave_frequency=OM, phase=0.0)
    with pytest.raises(sph2d.SphError, match="relaxation_init"):
        sim.wave_theory_update(w, 1)
    with pytest.raises(sph2d.SphError, match="zone"):
        sim.wave_maker_damping_zone(1)
    sim.relaxation_init(0.3)
    w.wave_type = sph2d.WAVE_SOLITARY
    with pytest.raises(sph2d.SphError, match="Stokes_2nd and Airy"):
        sim.wave_theory_update(w, 1)
    sim.close()
