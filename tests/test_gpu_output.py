"""GPU output path: Refresh_domain's vorticity and the step-start analytic fields
against the oracle, and the context-level writers (Output_for_tecplot.f90,
Output_for_paraview.f90, the restart record) against the downloaded state."""
import os
import struct
import sys

import numpy as np
import pytest

import sph2d
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_pair(n, steps, sample_every, precision=sph2d.FP64, integrator=2):
    prm = sph2d.tgv_params(n, integrator=integrator, sample_every=sample_every, precision=precision)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    o = Oracle(nx=n, ny=n, integrator=integrator, dt=prm.dt)
    o.init_tgv()
    for it in range(1, steps + 1):
        sim.time_step(it)
        o.time_step(it)
    return sim, o


def oracle_global(o, name, n):
    """Local-order oracle field in global (gid) order, fluid particles only."""
    loc = o.f64(name)
    g = o.i32("name")
    out = np.full(n, np.nan)
    sel = g < n
    out[g[sel]] = loc[sel]
    return out


@pytest.mark.parametrize("integrator,steps", [(2, 5), (2, 20), (3, 5)])
def test_vorticity_and_analytic_fields_match_the_oracle(integrator, steps):
    """Step 20 is an MLS step (its corrector is k_stage's); RK keeps the third
    evaluation's pairs. FP64 bar: 1e-12 of the field's largest magnitude."""
    say(f"test_vorticity_and_analytic_fields_match_the_oracle integrator={integrator} steps={steps}")
    n = 48
    sim, o = run_pair(n, steps, 5, integrator=integrator)
    w, ana, step = sim.output_fields()
    assert step == steps
    ref = oracle_global(o, "vorticity", n * n)
    err = np.max(np.abs(w - ref)) / np.max(np.abs(ref))
    say(f"vorticity rel err {err:.3e}")
    assert err < 1e-12
    for k, f in enumerate(("ana_u", "ana_v", "ana_p", "ana_vm")):
        r = o.f64(f)[: n * n]
        assert np.max(np.abs(ana[:, k] - r)) <= 1e-14 * np.max(np.abs(r)), f
    sim.close()


def test_vorticity_mixed_form_within_its_tolerance():
    say("test_vorticity_mixed_form_within_its_tolerance")
    n = 48
    sim, o = run_pair(n, 10, 10, precision=sph2d.FP32_MIXED)
    w, _, _ = sim.output_fields()
    ref = oracle_global(o, "vorticity", n * n)
    err = np.max(np.abs(w - ref)) / np.max(np.abs(ref))
    say(f"mixed vorticity rel err {err:.3e}")
    assert err < 1e-4
    sim.close()


def test_output_fields_need_a_sampling_step():
    say("test_output_fields_need_a_sampling_step")
    prm = sph2d.tgv_params(32, sample_every=10)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    sim.run(1, 3)
    with pytest.raises(sph2d.SphError) as e:
        sim.output_fields()
    assert e.value.code == 4
    sim.run(4, 7)
    with pytest.raises(sph2d.SphError):
        sim.output_for_tecplot(9, video_path=os.devnull)        # 10 was the last step
    sim.close()


def test_tecplot_and_paraview_files_of_a_context(tmp_path):
    say("test_tecplot_and_paraview_files_of_a_context")
    n = 40
    prm = sph2d.tgv_params(n, sample_every=10)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    sim.run(1, 10)
    video, surf, binf = tmp_path / "video_result.dat", tmp_path / "Free_Surface.dat", tmp_path / "b.dat"
    sim.output_for_tecplot(10, video, surf, binf)
    sim.output_for_paraview(10, tmp_path / "flow.vtu")
    st = sim.download_state()
    w, ana, _ = sim.output_fields()
    sc = sim.scales()
    assert (sc.domain_size_y, sc.velocity_0, sc.pressure_0, sc.time_0) == (1.0, 1.0, 1000.0, 1.0)
    ref = tmp_path / "ref.dat"
    sph2d.write_tecplot_zone(ref, 10 * prm.dt, st["x"], st["y"], st["u"], st["v"], st["p"], w, ana, sc)
    assert video.read_bytes() == ref.read_bytes()
    lines = video.read_text().split("\n")
    assert lines[2] == ' ZONE T=" %s ", I=%12d  F=POINT' % (sph2d.time_name(10 * prm.dt, 12, 5), n * n)
    row0 = [float(lines[3][20 * k:20 * k + 20]) for k in range(11)]
    assert abs(row0[6] - w[0]) < 1e-8 and abs(row0[0] - st["x"][0]) < 1e-8
    assert surf.read_text().split("\n")[2].endswith("I=           0  F=POINT")   # no surface in the TGV
    b = binf.read_bytes()
    assert b[:8] == b"#!TDV101"
    data = np.frombuffer(b[-48 * n * n:], dtype="<f8").reshape(n * n, 6)
    assert np.array_equal(data[:, 0], st["x"]) and np.array_equal(data[:, 4], st["p"] / 1000.0)
    vtu = (tmp_path / "flow.vtu").read_text()
    assert 'NumberOfPoints="%12d "' % (n * n) in vtu
    sim.close()


def test_restart_from_the_latest_results_file(tmp_path):
    """Write the restart zone at step 10, read it back (Distribute_Latest_Particle),
    continue in a new context: the state is the F20.8-rounded one, and the run
    stays within the restart's perturbation of the uninterrupted run (the restarted
    first predictor has no previous rates, as in the reference)."""
    say("test_restart_from_the_latest_results_file")
    n = 40
    prm = sph2d.tgv_params(n, sample_every=10)
    a = sph2d.Simulation(prm)
    a.init_tgv()
    a.run(1, 10)
    path = tmp_path / "Latest_Time_Results.dat"
    a.output_latest_results(10, path)
    st = a.download_state()
    a.run(11, 10)
    ke_a = a.sample()[0]
    r = sph2d.read_latest_results(path)
    assert r["step"] == 10 and len(r["x"]) == n * n
    assert np.max(np.abs(r["x"] - st["x"])) <= 5e-9 and np.max(np.abs(r["rho"] - st["rho"])) <= 5e-9
    b = sph2d.Simulation(prm)
    b.upload_state(r["x"], r["y"], r["u"], r["v"], r["rho"], r["mass"], r["energy"], r["type"])
    b.run(11, 10)
    s = b.sample()
    say(f"KE/KE0 uninterrupted {ke_a:.10f} restarted {s[0]:.10f}")
    assert np.all(np.isfinite(s))
    # KE0 of the restarted context is its own step-11 state: compare the kinetic energies
    sa, sb = a.download_state(), b.download_state()
    ea = np.sum(sa["mass"] * (sa["u"] ** 2 + sa["v"] ** 2))
    eb = np.sum(sb["mass"] * (sb["u"] ** 2 + sb["v"] ** 2))
    # the restarted first step's corrector adds only half an acceleration (zero
    # previous rates): ~0.5 dt |a| / U ~ 4e-4 in v, ~1e-3 in the kinetic energy
    assert abs(eb - ea) / ea < 5e-3
    a.close()
    b.close()


# --- Calculation_For_PostProceeding on the GPU ---------------------------------------------
GRID_TOL = 1e-10     # order-6 MLS: Cholesky vs the reference's Gauss-Jordan and another sum order


def _grid_cmp(g, ref, name):
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(g - ref)) / scale
    say(f"grid {name}: rel err {err:.3e}")
    assert err < GRID_TOL, name


def test_grid_postprocess_matches_the_oracle():
    say("test_grid_postprocess_matches_the_oracle")
    n = 48
    sim, o = run_pair(n, 10, 10)
    nex, ney, edx = sim.grid_size()
    assert (nex, ney) == (n, n) and edx == 1.0 / n
    g = sim.calculation_for_postproceeding(10)
    ref = o.grid_postprocess(10, nex, ney, edx)
    for col, name in enumerate(("u", "v", "p", "vorticity")):
        _grid_cmp(g[name], ref[:, col], name)
    assert np.all(g["fai"] == 1.0) and np.all(ref[:, 4] == 1.0)
    for col, name in ((5, "ana_u"), (6, "ana_v"), (7, "ana_p"), (8, "ana_vm")):
        assert np.max(np.abs(g[name] - ref[:, col])) <= 1e-14 * np.max(np.abs(ref[:, col])), name
    sim.close()


def test_grid_mls_is_exact_for_quadratic_velocity():
    """Per-routine path on an uploaded state whose velocity is a quadratic of position:
    interior nodes get the polynomial's value (order-6 MLS reproduces quadratics)."""
    say("test_grid_mls_is_exact_for_quadratic_velocity")
    n = 40
    o0 = Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    rng = np.random.default_rng(12)
    x = o0.f64("g_x")[:N] + rng.uniform(-0.2, 0.2, N) / n
    y = o0.f64("g_y")[:N] + rng.uniform(-0.2, 0.2, N) / n
    x, y = np.mod(x, 1.0), np.mod(y, 1.0)
    quad = lambda X, Y, c: c[0] + c[1] * X + c[2] * Y + c[3] * X * X + c[4] * X * Y + c[5] * Y * Y
    cu, cv = (0.3, 1.1, -0.7, 0.9, -1.3, 0.5), (-0.2, 0.4, 0.8, -0.6, 0.7, 1.2)
    prm = sph2d.tgv_params(n, sample_every=1)
    sim = sph2d.Simulation(prm)
    sim.upload_state(x, y, quad(x, y, cu), quad(x, y, cv), o0.f64("g_rho")[:N], o0.f64("g_mass")[:N])
    sim.distribute_period_boundary()
    sim.devide_block(1)
    sim.single_step_compute(1)
    sim.refresh_domain(1)
    g = sim.calculation_for_postproceeding(1)
    ex = np.tile(np.arange(n + 1) / n, n + 1)
    ey = np.repeat(np.arange(n + 1) / n, n + 1)
    inner = (ex > 0.05) & (ex < 0.95) & (ey > 0.05) & (ey < 0.95)
    for name, c in (("u", cu), ("v", cv)):
        err = np.abs(g[name] - quad(ex, ey, c))[inner].max()
        say(f"quadratic {name}: max err {err:.3e}")
        assert err < 1e-11
    sim.close()


def test_grid_level_set_on_a_free_surface_patch():
    """A fluid disc with an open boundary: surface particles and normals from the
    evaluation give the level set Fai and the nodes left empty outside; the MLS
    values inside match the oracle."""
    say("test_grid_level_set_on_a_free_surface_patch")
    n = 60
    o0 = Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    x, y = o0.f64("g_x")[:N], o0.f64("g_y")[:N]
    keep = ((x - 0.5) ** 2 + (y - 0.5) ** 2) < 0.3 ** 2
    rng = np.random.default_rng(9)
    st = (x[keep] + rng.uniform(-0.1, 0.1, keep.sum()) / n, y[keep] + rng.uniform(-0.1, 0.1, keep.sum()) / n,
          o0.f64("g_u")[:N][keep], o0.f64("g_v")[:N][keep], o0.f64("g_rho")[:N][keep], o0.f64("g_mass")[:N][keep])
    prm = sph2d.tgv_params(n, sample_every=1)
    sim = sph2d.Simulation(prm, capacity=N)
    sim.upload_state(*st)
    o = Oracle(nx=n, ny=n, dt=prm.dt)
    o.set_fluid_state(*st)
    sim.distribute_period_boundary()
    sim.devide_block(1)
    sim.single_step_compute(1)
    sim.refresh_domain(1)
    o.period_boundary()
    o.devide_block(1)
    o.single_step_compute(1)
    o.refresh()
    nex, ney, edx = sim.grid_size()
    g = sim.calculation_for_postproceeding(1)
    ref = o.grid_postprocess(1, nex, ney, edx)
    say(f"surface particles {int(o.i32('g_surface').sum())}, empty nodes {int((ref[:, 4] <= -0.9).sum())}")
    assert (ref[:, 4] <= -0.9).sum() > 100 and (ref[:, 4] == 1.0).sum() > 100
    assert np.max(np.abs(g["fai"] - ref[:, 4])) < 1e-12
    for col, name in enumerate(("u", "v", "p")):
        _grid_cmp(g[name], ref[:, col], name)
    sim.close()


def test_grid_output_files_of_a_context(tmp_path):
    say("test_grid_output_files_of_a_context")
    n = 32
    prm = sph2d.tgv_params(n, sample_every=10)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    sim.run(1, 10)
    grid, binf = tmp_path / "Post_Proceeding_Grid.dat", tmp_path / "grid.dat"
    sim.output_for_grid_postproceding(10, grid, binf)
    g = sim.calculation_for_postproceeding(10)
    nex, ney, edx = sim.grid_size()
    ref = tmp_path / "ref.dat"
    ana = np.stack([g["ana_u"], g["ana_v"], g["ana_p"], g["ana_vm"]], axis=1)
    sph2d.write_grid_zone(ref, 10 * prm.dt, nex, ney, edx, g["u"], g["v"], g["p"], g["vorticity"], g["fai"], ana,
                          sim.scales())
    assert grid.read_bytes() == ref.read_bytes()
    nn = (nex + 1) * (ney + 1)
    vals = np.frombuffer(binf.read_bytes()[-56 * nn:], dtype="<f8").reshape(nn, 7)
    assert np.array_equal(vals[:, 2], g["u"]) and np.array_equal(vals[:, 6], g["fai"])
    sim.close()


# --- Sampling_subroutine on the GPU -----------------------------------------------------------
def test_sampling_points_match_the_oracle():
    say("test_sampling_points_match_the_oracle")
    n = 48
    sim, o = run_pair(n, 10, 10)
    N = n * n
    rng = np.random.default_rng(21)
    gx, gy = o.f64("g_x")[:N], o.f64("g_y")[:N]
    sx = np.concatenate([rng.uniform(0.0, 1.0, 300), gx[[3, 400, 2000]]])
    sy = np.concatenate([rng.uniform(0.0, 1.0, 300), gy[[3, 400, 2000]]])
    g = sim.sampling(10, sx, sy)
    ref = o.sampling_points(sx, sy)
    assert np.max(np.abs(g["average_p"] - ref[:, 0])) <= 1e-13 * np.max(np.abs(ref[:, 0]))
    for col, name in ((1, "mls_p"), (2, "u"), (3, "v")):
        _grid_cmp(g[name], ref[:, col], name)
    with pytest.raises(sph2d.SphError):
        sim.sampling(10, [5.0], [0.5])                  # outside the background grid
    with pytest.raises(sph2d.SphError):
        sim.sampling(9, sx, sy)                          # not the last sampling step
    sim.close()


def test_sampling_and_wave_probes_on_a_fluid_disc():
    """Probes around the rim of a fluid disc: zeros with 3 or fewer particles in the
    support, Shepard values with 4 to 10, MLS inside, as the oracle; the wave front
    and the gauge elevations exactly as the oracle."""
    say("test_sampling_and_wave_probes_on_a_fluid_disc")
    n = 40
    o0 = Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    x, y = o0.f64("g_x")[:N], o0.f64("g_y")[:N]
    keep = ((x - 0.5) ** 2 + (y - 0.5) ** 2) < 0.3 ** 2
    st = [a[keep] for a in (x, y, o0.f64("g_u")[:N], o0.f64("g_v")[:N], o0.f64("g_rho")[:N],
                            o0.f64("g_mass")[:N])]
    prm = sph2d.tgv_params(n, sample_every=1)
    sim = sph2d.Simulation(prm, capacity=N)
    sim.upload_state(*st)
    o = Oracle(nx=n, ny=n, dt=prm.dt)
    o.set_fluid_state(*st)
    for s in (sim,):
        s.distribute_period_boundary()
        s.devide_block(1)
        s.single_step_compute(1)
        s.refresh_domain(1)
    o.period_boundary()
    o.devide_block(1)
    o.single_step_compute(1)
    o.refresh()
    ang = np.linspace(0, 2 * np.pi, 200, endpoint=False)
    rad = np.concatenate([[0.0, 0.1, 0.2], np.linspace(0.27, 0.42, 10)])
    sx = (0.5 + np.outer(rad, np.cos(ang))).ravel()
    sy = (0.5 + np.outer(rad, np.sin(ang))).ravel()
    g = sim.sampling(1, sx, sy, gravity=9.81)
    ref = o.sampling_points(sx, sy, 9.81)
    zero = np.all(ref == 0.0, axis=1)
    say(f"probes {len(sx)}, empty {int(zero.sum())}")
    assert zero.sum() > 100 and (~zero).sum() > 1000
    for col, name in enumerate(("average_p", "mls_p", "u", "v")):
        assert np.all(g[name][zero] == 0.0), name
    assert np.max(np.abs(g["average_p"] - ref[:, 0])) <= 1e-13 * np.max(np.abs(ref[:, 0]))
    for col, name in ((1, "mls_p"), (2, "u"), (3, "v")):
        _grid_cmp(g[name], ref[:, col], name)
    px = np.array([0.2, 0.21, 0.5, 0.79, 0.95, -0.5])
    front, el = sim.wave_probes(px)
    rf, rel = o.wave_probes(px)
    say(f"front {front!r} oracle {rf!r}")
    assert front == rf and np.array_equal(el, rel)
    assert el[4] == 0.0 and el[5] == 0.0 and el[2] > 0.75
    sim.close()
    o.close()


def test_sampling_files_of_a_context(tmp_path):
    """The sampling-line result files (Sampling_subroutine.f90:466-512) and the probe
    files written by sph2d.SamplingProbes: values as the oracle's at the same points,
    to the 6 (20f12.6) and 8 (20F20.8) decimals written."""
    say("test_sampling_files_of_a_context")
    n = 40
    sim, o = run_pair(n, 10, 10)
    first = np.array([[0.0, 0.25], [0.0, 0.5], [0.0, 0.75]])
    last = np.array([[1.0, 0.25], [1.0, 0.5], [1.0, 0.75]])
    sim.output_sampling_lines(10, tmp_path, first, last, 41)
    sc = sim.scales()
    tn = 10 * sim.params.dt / (1.0 / sc.velocity_0)
    for k in range(3):
        name = tmp_path / f"{k + 1}th_{sph2d.time_name(tn, 7, 4)}_SamplingLine_Results.dat"
        rows = np.loadtxt(name)
        assert rows.shape == (41, 9)
        x, y = sph2d.sampling_line_points(first[k], last[k], 41)
        ref = o.sampling_points(x, y)
        np.testing.assert_allclose(rows[:, 0], x, atol=6e-7)
        np.testing.assert_allclose(rows[:, 1], y, atol=6e-7)
        np.testing.assert_allclose(rows[:, 5], ref[:, 2] / sc.velocity_0, atol=1.1e-6)
        np.testing.assert_allclose(rows[:, 6], ref[:, 3] / sc.velocity_0, atol=1.1e-6)
        np.testing.assert_allclose(rows[:, 7], ref[:, 1] / sc.pressure_0, atol=1.1e-6)
        np.testing.assert_allclose(rows[:, 8], ref[:, 0] / sc.pressure_0, atol=1.1e-6)
        np.testing.assert_allclose(rows[:, 2], np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) *
                                   np.exp(-8 * np.pi ** 2 * sim.params.nu / sim.params.U * tn), atol=1.1e-6)
    pts = [(0.3, 0.3), (0.55, 0.8)]
    probes = sph2d.SamplingProbes(sim, tmp_path / "Sampling", points=pts, probe_x=[0.5],
                                  sampling_time_step=10, wave_front_dir=tmp_path)
    probes(10)
    ref = o.sampling_points([0.3, 0.55], [0.3, 0.8])
    front, _ = o.wave_probes([0.5])
    t = 10 * sim.params.dt
    line = (tmp_path / "Wave_front_results.dat").read_text()
    assert line == " " + sph2d.format_list_directed(t * sc.time_0) + " " + \
        sph2d.format_list_directed(front / sc.domain_size_y) + "\n"
    for m in range(2):
        v = np.loadtxt(tmp_path / "Sampling" / "Press" / f"Probe_{m + 1}_Normalized_Pressure.dat")
        np.testing.assert_allclose(v, [t * sc.time_0, ref[m, 1] / sc.pressure_0], atol=1.1e-8)
        kal = np.loadtxt(tmp_path / "Sampling" / "Press" / f"Probe_{m + 1}_Kalman_Pressure.dat")
        assert np.array_equal(kal, v)                    # the first sampling passes the value through
    assert (tmp_path / "Sampling" / "WaveHeight" / "wave_gauge_1_result.dat").exists()
    sim.close()


def test_wavemaker_pull_back_matches_the_oracle():
    say("test_wavemaker_pull_back_matches_the_oracle")
    n = 40
    prm = sph2d.tgv_params(n)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    g0 = sim.download_state()
    o = Oracle(nx=n, ny=n, dt=prm.dt)
    o.set_fluid_state(g0["x"], g0["y"], g0["u"], g0["v"], g0["rho"], g0["mass"])
    xp, up = 0.3 + 0.37 / n, 0.25
    sim.wavemaker_pull_back(xp, up)
    o.wavemaker_pull_back(xp, up)
    g = sim.download_state()
    N = n * n
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v")):
        assert np.array_equal(g[f], o.f64(of)[:N]), f
    assert np.all(g["x"] >= xp)
    sim.close()
