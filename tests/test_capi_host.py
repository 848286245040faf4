"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, the Python mirror's struct layouts match the C header, and with no
GPU it fails loudly (no CPU fallback)."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import sph2d
from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "sph2d.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sph2d_\w+)\s*\(", text, re.M)))


def test_header_and_python_mirror_list_the_same_entry_points():
    assert header_functions() == sorted(sph2d.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(sph2d.LIB_PATH):
        pytest.skip("libsph2d.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", sph2d.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (sph2d_\w+)", out))
    missing = set(header_functions()) - exported
    assert not missing, missing
    lib = ctypes.CDLL(sph2d.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name)


def _c_layout(struct, fields):
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           f'printf("%zu\\n", sizeof({struct}));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof({struct}, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c11", "-o", exe, c], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    return vals[0], vals[1:]


@pytest.mark.parametrize("py,cname", [(sph2d.Params, "sph2d_params"), (sph2d.Diag, "sph2d_diag"),
                                     (sph2d.Scales, "sph2d_scales")])
def test_struct_layout_matches_header(py, cname):
    names = [f[0] for f in py._fields_]
    size, offs = _c_layout(cname, names)
    assert ctypes.sizeof(py) == size
    assert [getattr(py, n).offset for n in names] == offs


def test_no_gpu_fails_loudly():
    if gpu_available():
        pytest.skip("a GPU is present")
    if not os.path.exists(sph2d.LIB_PATH):
        pytest.skip("libsph2d.so not built")
    p = sph2d.default_params()
    assert p.struct_size == ctypes.sizeof(sph2d.Params)
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code in (3,)
    with pytest.raises(sph2d.SphError):
        sph2d.status_equation([1000.0])


def test_bad_parameters_are_rejected_before_any_device_call():
    p = sph2d.default_params(integrator=7)
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code == 5
    p = sph2d.default_params(dt=-1.0)
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code == 1
    p = sph2d.default_params()
    p.struct_size = 8
    with pytest.raises(sph2d.SphError) as e:
        sph2d.Simulation(p)
    assert e.value.code == 1


def test_tgv_params_scale_with_lattice():
    p = sph2d.tgv_params(2048)
    assert p.nx == p.ny == 2048
    assert p.dx == 1.0 / 2048 and p.h == 1.2 * p.dx and p.max_shift == 0.25 * p.dx
    assert p.dt == pytest.approx(1e-4 * 100 / 2048)


def test_fortran_driver_links_and_fails_loudly_without_gpu():
    exe = os.path.join(os.path.dirname(sph2d.LIB_PATH), "fortran", "tgv_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (flang absent)")
    if gpu_available():
        pytest.skip("a GPU is present")
    out = subprocess.run([exe, "16", "1"], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "runs on the GPU only" in out.stdout


# --- piston wavemaker kinematics (host arithmetic, no GPU) -----------------------------------
def _wm(kind, **kw):
    w = sph2d.Wavemaker(wave_type=kind, soft_steps=1000, standing_steps=0, dt=1e-4, g=9.81, water_depth=0.4,
                        wave_height=0.05, wave_number=2 * np.pi / 2.0, wave_frequency=2 * np.pi / 1.5,
                        phase=np.pi / 2)
    for k, v in kw.items():
        setattr(w, k, v)
    return w


@pytest.mark.parametrize("kind", [sph2d.WAVE_AIRY, sph2d.WAVE_STOKES2])
def test_piston_velocity_and_acceleration_are_the_time_derivatives(kind):
    """Wavemaker_Subroutine.f90:93-135 after the soft start: central differences of
    the position over one step give the velocity, and of the velocity the
    acceleration (the laws are smooth; the step is 1e-4 s)."""
    w = _wm(kind)
    for it in (1500, 4321, 9000):
        xm, um, _ = sph2d.wavemaker_motion(w, it - 1)
        x0, u0, a0 = sph2d.wavemaker_motion(w, it)
        xp, up, _ = sph2d.wavemaker_motion(w, it + 1)
        assert abs((xp - xm) / (2 * w.dt) - u0) < 1e-6 * max(abs(u0), 1e-3)
        assert abs((up - um) / (2 * w.dt) - a0) < 1e-5 * max(abs(a0), 1e-2)


def test_piston_stroke_is_biesel_and_soft_start_is_a_step():
    """The first-order stroke is the textbook piston transfer function
    S/H = (sinh 2kd + 2kd) / (2 (cosh 2kd - 1)) (written differently from the code's
    m1), and the soft-start coefficient divides integers (:69-73): the paddle stays
    at rest before step soft_steps and moves with full amplitude from it."""
    w = _wm(sph2d.WAVE_AIRY)
    k, d, H = w.wave_number, w.water_depth, w.wave_height
    S = H * (np.sinh(2 * k * d) + 2 * k * d) / (2 * (np.cosh(2 * k * d) - 1))
    xs = [sph2d.wavemaker_motion(w, it)[0] for it in range(1000, 1000 + 15000, 7)]
    assert abs(max(np.abs(xs)) - S / 2) < 1e-6 * S
    assert all(sph2d.wavemaker_motion(w, it) == (0.0, 0.0, 0.0) for it in (0, 1, 500, 999))
    x, u, a = sph2d.wavemaker_motion(w, 1000)
    th = w.wave_frequency * 1000 * w.dt + w.phase
    assert abs(x - S / 2 * np.sin(th)) < 1e-12 and u != 0.0


def test_stokes_second_order_part_is_a_second_harmonic():
    """Stokes_2nd minus Airy is b sin(2 theta) with one constant b over time."""
    wa, ws = _wm(sph2d.WAVE_AIRY), _wm(sph2d.WAVE_STOKES2)
    bs = []
    for it in range(1003, 20000, 997):
        th = ws.wave_frequency * it * ws.dt + ws.phase
        if abs(np.sin(2 * th)) < 0.2:
            continue
        bs.append((sph2d.wavemaker_motion(ws, it)[0] - sph2d.wavemaker_motion(wa, it)[0]) / np.sin(2 * th))
    assert len(bs) > 5 and np.ptp(bs) < 1e-9 * max(1.0, abs(np.mean(bs)))


def test_solitary_paddle_follows_goring():
    """wavemaker_plate_motion.f90 (Goring 1978): the position solves
    X = H/K (tanh(K/d (c t - X - lambda)) + tanh(K lambda/d)) to the iteration's
    tolerance, the full stroke is H/K (1 + tanh(acosh(1/sqrt(ez)))) = H/K (1 + sqrt(1 - ez)),
    just short of Goring's sqrt(16 H d / 3), the largest paddle speed is
    H c / (d + H), and the velocity is the derivative of the position."""
    H, d, g = 0.1, 0.4, 9.81
    w = _wm(sph2d.WAVE_SOLITARY, wave_height=H, water_depth=d, g=g, dt=1e-3)
    K = np.sqrt(3 * H / (4 * d))
    c = np.sqrt(g * (d + H))
    ez = float(np.float32(0.002))
    lam = np.log(1 / np.sqrt(ez) * (1 + np.sqrt(1 - ez))) / K * d
    ts, xs, us = [], [], []
    for it in range(1, 8001, 10):
        x, u, a = sph2d.wavemaker_motion(w, it)
        t = it * w.dt
        res = x - H / K * (np.tanh(K / d * (c * t - x - lam)) + np.tanh(K * lam / d))
        assert abs(res) < 2e-4
        ts.append(t); xs.append(x); us.append(u)
    xs, us = np.array(xs), np.array(us)
    assert abs(xs[-1] - H / K * (1 + np.sqrt(1 - ez))) < 2e-4
    assert abs(xs[-1] - np.sqrt(16 * H * d / 3)) < 1e-3
    assert abs(us.max() - H * c / (d + H)) < 1e-3 * H * c / (d + H)
    fd = np.gradient(xs, np.array(ts))
    assert np.max(np.abs(fd - us)) < 0.02 * us.max()


def test_moving_square_layout_matches_the_oracle():
    """sph2d_moving_square_layout (a host routine, no GPU) and the oracle's restatement of
    MS/Initial_Particle_From_First_Step.f90 give the same particles in the same order:
    positions, boundary types and initial types bit for bit."""
    if not os.path.exists(sph2d.LIB_PATH):
        pytest.skip("libsph2d.so not built")
    from oracle import pyoracle as po
    for (nx, ny, bx, by, cx, cy, L, layers) in ((100, 48, 2.0, 0.96, 0.6, 0.48, 0.4, 3),
                                                (200, 100, 10.0, 5.0, 2.5, 2.5, 1.0, 4)):
        p = sph2d.moving_square_params(n_per_d=int(round(nx / bx)), tank=(bx, by), layers=layers,
                                       nx=nx, ny=ny, body_cx=cx, body_cy=cy, body_len=L)
        x, y, bt, it, nf = sph2d.moving_square_layout(p)
        o = po.Oracle(ms=1, nx=nx, ny=ny, box_x=bx, box_y=by, body_cx=cx, body_cy=cy, body_len=L, walls=layers,
                      h_factor=1.15)
        o.init_moving_square()
        n = o.n_global
        assert nf == o.n_fluid and len(x) == n
        assert np.array_equal(x, o.f64("g_x")[:n]) and np.array_equal(y, o.f64("g_y")[:n])
        assert np.array_equal(it, o.i32("g_initial_type")[:n])
        assert np.array_equal(bt[nf:], o.i32("g_btype")[nf:n])
