"""Pins the oracle functions that no compiled reference piece covers: the free
surface (Define_freesurface_and_Normailized_derivates.f90:65-440), particle
shifting (Shift_particle_position.f90:56-136), the predictor-corrector corrector
(time_iteration.f90:541-657) and the single-rank local order
(Devide_block_and_send_tasks.f90:75-83,123-162,488-633).

Each check is chosen so that a plausible mistake in the oracle fails it:
* free surface: an O(N^2) numpy brute force of the volume fraction, the kernel
  gradient sums, the surface set, the normals, the geometric un-flag rule and the
  near-surface set on a perturbed, holed disc; normals against the outward radial
  direction; hand-built clusters where each geometric rule does or does not fire;
* shift: zero on the perfect periodic lattice; the two-particle closed form from
  the analytic derivative of the truncated Gaussian, with the (W/W_dx)^4 factor
  exactly 1 + R at r = dx; superposition of three particles; the 0.25 dx clamp,
  the density Taylor correction and the EOS on an interior state;
* corrector: a lone particle under gravity follows the closed form of the
  reference's trapezoid step, including its extra dt^2 a / 2 position term;
* local order: a numpy lexsort over (cell column, cell row, gid) with the
  reference's cell index floor((x - x0)/dx_c + 1), ill particles typed 120, and a
  hand-worked four-particle case.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

PI_MODULE = 3.14159265358979324            # information_module.f90:29
PI4 = float(np.float32(3.14159265358))     # compute_kernel.f90:32 (REAL(4) literal)
E9 = float(np.float32(np.exp(np.float32(-9.0))))
C4 = float(np.float32(np.float32(1.0) - np.float32(10.0) * np.float32(E9)))
RMIN = float(np.float32(1.0e-7))
F065 = float(np.float32(0.65))
F1414 = float(np.float32(1.414))


def kernel(pdx, pdy, hb):
    """Improved Gaussian (compute_kernel.f90:123-133), vectorised, with the REAL(4)
    constants; pinned bit for bit against the compiled reference elsewhere."""
    r = np.sqrt(pdx * pdx + pdy * pdy)
    q = r / hb
    factor = 1.0 / (PI4 * (hb * hb) * C4)
    w = (np.exp(-(q * q)) - E9) * factor
    g = np.exp(-(q * q)) * (-2 * q) / (r * hb) * factor
    return w, g * pdx, g * pdy


def test_numpy_kernel_equals_oracle_kernel():
    rng = np.random.default_rng(0)
    for _ in range(200):
        h = 0.012
        r = rng.uniform(1e-6, 3 * h)
        th = rng.uniform(0, 2 * np.pi)
        pd = np.array([r * np.cos(th), r * np.sin(th)])
        r = float(np.sqrt(pd[0] * pd[0] + pd[1] * pd[1]))
        w0, d0 = po.compute_kernel(r, pd, h)
        w1, dx1, dy1 = kernel(np.array([pd[0]]), np.array([pd[1]]), h)
        assert abs(w1[0] - w0) <= 1e-15 * abs(w0) + 1e-300
        assert abs(dx1[0] - d0[0]) <= 1e-15 * abs(d0[0]) + 1e-300
        assert abs(dy1[0] - d0[1]) <= 1e-15 * abs(d0[1]) + 1e-300


def run_eval(o, periodic=True):
    if periodic:
        o.period_boundary()
    o.devide_block(1)
    o.single_step_compute(1)


def brute_pairs(x, y, typ, h):
    """All unordered pairs of find_particle_pair.f90:64-123 (one of them fluid,
    1e-7 <= r <= 3 hbar), by brute force over the local arrays."""
    n = len(x)
    i, j = np.triu_indices(n, 1)
    pdx, pdy = x[i] - x[j], y[i] - y[j]
    r = np.sqrt(pdx * pdx + pdy * pdy)
    keep = ((typ[i] == 2) | (typ[j] == 2)) & (r <= 3 * h) & (r >= RMIN)
    return i[keep], j[keep], pdx[keep], pdy[keep]


def freesurface_brute(o):
    """Brute-force restatement of Define_freesurface_and_Normailized_derivates.f90:
    fra_v (:79-93), surface candidates (:119-131), sum dW (:141-159), normals and
    the geometric un-flag tests (:295-376), near-surface flags (:418-440)."""
    x, y = o.f64("x"), o.f64("y")
    typ, vol = o.i32("type"), o.f64("volume")
    h = float(o.f64("h")[0])
    n = len(x)
    i, j, pdx, pdy = brute_pairs(x, y, typ, h)
    sv = PI_MODULE * ((3 * h) * (3 * h))
    fra = np.zeros(n)
    np.add.at(fra, i, vol[j] / sv)
    np.add.at(fra, j, vol[i] / sv)
    _, gx, gy = kernel(pdx, pdy, h)
    sdx, sdy = np.zeros(n), np.zeros(n)
    np.add.at(sdx, i, gx * vol[j]); np.add.at(sdy, i, gy * vol[j])
    np.add.at(sdx, j, -gx * vol[i]); np.add.at(sdy, j, -gy * vol[i])
    cand = (typ == 2) & (fra <= F065)
    nrm = np.full((n, 2), np.nan)
    surf = cand.copy()
    fluid = np.flatnonzero(typ == 2)
    for s in np.flatnonzero(cand):
        ln = np.sqrt(sdx[s] ** 2 + sdy[s] ** 2)
        if not np.isfinite(ln) or ln == 0.0:
            surf[s] = False
            continue
        nx, ny = -sdx[s] / ln, -sdy[s] / ln
        nrm[s] = nx, ny
        others = fluid[fluid != s]            # self: acos(0/0) is NaN, never <= pi/4
        qx, qy = x[others] - x[s], y[others] - y[s]
        ji = np.sqrt(qx * qx + qy * qy)
        nd = nx * qx + ny * qy
        tx, ty = x[others] - (x[s] + h * nx), y[others] - (y[s] + h * ny)
        jt = np.sqrt(tx * tx + ty * ty)
        lim = F1414 * h
        with np.errstate(invalid="ignore"):
            rule1 = (ji >= lim) & (jt < h)
            rule2 = (ji < lim) & (np.arccos(nd / ji) <= PI_MODULE / 4.0)
        if np.any(rule1 | rule2):
            surf[s] = False
    near = np.zeros(n, dtype=bool)
    near[j[surf[i]]] = True
    near[i[surf[j]]] = True
    return dict(fra=fra, sdx=sdx, sdy=sdy, cand=cand, nrm=nrm, surf=surf, near=near)


def holed_disc(n=60, R=0.3, amp=0.3, drop=0.08, seed=9):
    o0 = po.Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    x, y = o0.f64("g_x")[:N], o0.f64("g_y")[:N]
    u, v = o0.f64("g_u")[:N], o0.f64("g_v")[:N]
    rho, m = o0.f64("g_rho")[:N], o0.f64("g_mass")[:N]
    o0.close()
    rng = np.random.default_rng(seed)
    keep = (((x - 0.5) ** 2 + (y - 0.5) ** 2) < R * R) & (rng.uniform(size=N) > drop)
    k = int(keep.sum())
    xs = x[keep] + rng.uniform(-amp, amp, k) / n
    ys = y[keep] + rng.uniform(-amp, amp, k) / n
    return xs, ys, u[keep], v[keep], rho[keep], m[keep]


def test_free_surface_brute_force_on_a_holed_disc():
    n = 60
    st = holed_disc(n)
    o = po.Oracle(nx=n, ny=n, integrator=2)
    o.set_fluid_state(*st)
    run_eval(o)
    b = freesurface_brute(o)
    fra = o.f64("fra_v")
    np.testing.assert_allclose(fra, b["fra"], rtol=1e-13, atol=0)
    # no particle sits within rounding of the 0.65 threshold, so the sets are exact
    assert np.min(np.abs(b["fra"] - F065)) > 1e-9
    np.testing.assert_allclose(o.f64("sum_dwdx_x"), b["sdx"], rtol=0, atol=1e-12 * np.max(np.abs(b["sdx"])))
    np.testing.assert_allclose(o.f64("sum_dwdx_y"), b["sdy"], rtol=0, atol=1e-12 * np.max(np.abs(b["sdy"])))
    cand = b["cand"]
    nrm_o = np.stack([o.f64("normal_x"), o.f64("normal_y")], axis=1)
    np.testing.assert_allclose(nrm_o[cand], b["nrm"][cand], rtol=0, atol=1e-12)
    surf_o = o.i32("surface").astype(bool)
    assert np.array_equal(surf_o, b["surf"])
    assert np.array_equal(o.i32("near_surface").astype(bool), b["near"])
    # the geometric tests act: some candidates (holes inside the disc) are un-flagged
    assert cand.sum() > surf_o.sum() + 5, (cand.sum(), surf_o.sum())
    assert surf_o.sum() > 60
    # physics: the kept surface lies on the rim, with outward unit normals
    x, y = o.f64("x"), o.f64("y")
    rx, ry = x - 0.5, y - 0.5
    rad = np.sqrt(rx * rx + ry * ry)
    assert np.all(rad[surf_o] > 0.3 - 3.0 / n)
    cosang = (nrm_o[surf_o, 0] * rx[surf_o] + nrm_o[surf_o, 1] * ry[surf_o]) / rad[surf_o]
    assert np.median(cosang) > 0.95 and np.min(cosang) > 0.5
    # near-surface particles are exactly the partners of surface particles, all within 3h
    assert np.all(o.i32("near_surface")[surf_o] >= 0)


def cluster(extra, n=40):
    """A half-disc of lattice particles left of particle i = (0.5, 0.5) (its
    outward normal is +x) plus extra particles at offsets (in units of h) from i.
    Returns the oracle after one evaluation and i's local index."""
    dx = 1.0 / n
    h = 1.2 * dx
    pts = [(0.5, 0.5)]                          # gid 0 is particle i
    for a in range(0, 7):
        for b in range(-7, 8):
            if a * a + b * b <= 49 and (a, b) != (0, 0):
                pts.append((0.5 - a * dx, 0.5 + b * dx))
    for ex, ey in extra:
        pts.append((0.5 + ex * h, 0.5 + ey * h))
    p = np.array(pts)
    k = len(p)
    rho = np.full(k, 1000.0)
    o = po.Oracle(nx=n, ny=n, integrator=2)
    o.set_fluid_state(p[:, 0], p[:, 1], np.zeros(k), np.zeros(k), rho, rho * dx * dx)
    run_eval(o)
    name = o.i32("name")
    li = int(np.flatnonzero(name == 0)[0])     # gid 0 is (0.5, 0.5)
    return o, li


@pytest.mark.parametrize("extra,expect", [
    ([], 1),                       # open half-plane: stays a surface particle
    ([(1.5, 0.0)], 0),             # JI = 1.5h >= 1.414h and JT = 0.5h < h: un-flagged (:356)
    ([(1.0, 0.3)], 0),             # JI < 1.414h at 16.7 deg from the normal: un-flagged (:363)
    ([(0.5, 1.2)], 1),             # JI = 1.3h at 67 deg: no rule fires
    ([(2.5, 0.0)], 1),             # JI = 2.5h but JT = 1.5h > h: no rule fires
    ([(1.2, 1.2)], 1),             # JI = 1.70h, JT = 1.22h: no rule fires
    ([(1.8, 0.3)], 0),             # JI = 1.82h, JT = 0.85h: un-flagged
])
def test_geometric_surface_rule_on_hand_built_clusters(extra, expect):
    o, li = cluster(extra)
    fra = o.f64("fra_v")[li]
    assert fra <= F065                          # i is a surface candidate in every case
    nx, ny = o.f64("normal_x")[li], o.f64("normal_y")[li]
    assert nx > 0.97                            # normal within ~14 deg of +x
    assert o.i32("surface")[li] == expect
    b = freesurface_brute(o)
    assert np.array_equal(o.i32("surface").astype(bool), b["surf"])
    assert np.array_equal(o.i32("near_surface").astype(bool), b["near"])


def cluster_tuned(fra_target, n=40):
    """The half-disc cluster plus an arc of particles at 2.6 h right of i (too far
    for either geometric rule: JT >= JI - h > h) and one more on the arc whose
    volume sets i's volume fraction to fra_target."""
    dx = 1.0 / n
    h = 1.2 * dx
    pts = [(0.5, 0.5)]
    for a in range(0, 7):
        for b in range(-7, 8):
            if a * a + b * b <= 49 and (a, b) != (0, 0):
                pts.append((0.5 - a * dx, 0.5 + b * dx))
    for th in np.linspace(-1.2, 1.2, 4):
        pts.append((0.5 + 2.6 * h * np.cos(th), 0.5 + 2.6 * h * np.sin(th)))
    # a small jitter breaks the exact 45-degree ties of the perfect lattice, where
    # the last bit of a normal decides the acos test (i itself stays at the centre)
    rng = np.random.default_rng(21)
    pts = [pts[0]] + [(px + rng.uniform(-0.002, 0.002) * dx, py + rng.uniform(-0.002, 0.002) * dx)
                      for px, py in pts[1:]]

    def run(extra_vol):
        p = list(pts)
        vol = [dx * dx] * len(p)
        if extra_vol is not None:
            p.append((0.5 + 2.6 * h * np.cos(1.4), 0.5 + 2.6 * h * np.sin(1.4)))
            vol.append(extra_vol)
        p, vol = np.array(p), np.array(vol)
        k = len(p)
        rho = np.full(k, 1000.0)
        o = po.Oracle(nx=n, ny=n, integrator=2)
        o.set_fluid_state(p[:, 0], p[:, 1], np.zeros(k), np.zeros(k), rho, rho * vol)
        run_eval(o)
        li = int(np.flatnonzero(o.i32("name") == 0)[0])
        return o, li

    o, li = run(None)
    sv = PI_MODULE * ((3 * h) * (3 * h))
    base = o.f64("fra_v")[li]
    assert base < fra_target
    return run((fra_target - base) * sv)


@pytest.mark.parametrize("side", [-1, 1])
def test_surface_volume_fraction_threshold(side):
    """Define_freesurface_and_Normailized_derivates.f90:123: a fluid particle is a
    surface candidate when fra_v <= 0.65 (REAL(4) literal); here nothing else
    decides, so i is a surface particle just below the threshold and not just above."""
    # 1e-8 either side: inside the 2.4e-8 gap between the REAL(4) 0.65 and 0.65d0
    o, li = cluster_tuned(F065 + side * 1e-8)
    fra = o.f64("fra_v")[li]
    assert abs(fra - (F065 + side * 1e-8)) < 1e-12
    assert o.i32("surface")[li] == (1 if side < 0 else 0)
    b = freesurface_brute(o)
    assert np.array_equal(o.i32("surface").astype(bool), b["surf"])


# --- Shift_particle_position -------------------------------------------------------
def w_analytic(r, h):
    factor = 1.0 / (PI4 * h * h * C4)
    return (np.exp(-(r / h) ** 2) - E9) * factor


def dwdr_analytic(r, h):
    """d/dr of the truncated Gaussian, written from the formula W(r) above."""
    factor = 1.0 / (PI4 * h * h * C4)
    return -2.0 * r / (h * h) * np.exp(-(r / h) ** 2) * factor


def shift_pair_term(xi, xj, vol_j, h, dx, beta=0.5, R=0.2):
    """delta r_i of one neighbour j: -beta h^2 (1 + R (W/W_dx)^4) grad_i W V_j."""
    d = xi - xj
    r = np.sqrt(d @ d)
    grad = dwdr_analytic(r, h) * d / r
    ratio = w_analytic(r, h) / w_analytic(dx, h)
    return -beta * h * h * (1 + R * ratio ** 4) * grad * vol_j


def small_state(pts, n=40):
    dx = 1.0 / n
    p = np.array(pts)
    k = len(p)
    rho = np.full(k, 1000.0)
    o = po.Oracle(nx=n, ny=n, integrator=2)
    o.set_fluid_state(p[:, 0], p[:, 1], np.zeros(k), np.zeros(k), rho, rho * dx * dx)
    run_eval(o)
    o.shift_sums()
    name = o.i32("name")
    inv = np.argsort(name)
    sh = np.stack([o.f64("shift_x"), o.f64("shift_y")], axis=1)[inv]
    vol = o.f64("volume")[inv]
    return sh, vol, dx, 1.2 * dx


@pytest.mark.parametrize("d_over_dx,theta", [(1.0, 0.0), (1.0, 0.9), (1.7, 2.2), (3.4, -0.4)])
def test_shift_two_particle_closed_form(d_over_dx, theta):
    n = 40
    dx = 1.0 / n
    d = d_over_dx * dx
    pi = np.array([0.5, 0.5])
    pj = pi + d * np.array([np.cos(theta), np.sin(theta)])
    sh, vol, dx, h = small_state([pi, pj], n)
    exp_i = shift_pair_term(pi, pj, vol[1], h, dx)
    exp_j = shift_pair_term(pj, pi, vol[0], h, dx)
    np.testing.assert_allclose(sh[0], exp_i, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(sh[1], exp_j, rtol=1e-13, atol=1e-300)
    # repulsive: each particle moves away from the other
    assert np.dot(sh[0], pi - pj) > 0 and np.dot(sh[1], pj - pi) > 0
    if d_over_dx == 1.0:
        # at r = dx the shape factor is exactly 1 + R
        g = dwdr_analytic(dx, h)
        assert np.linalg.norm(sh[0]) == pytest.approx(0.5 * h * h * 1.2 * abs(g) * vol[1], rel=1e-13)


def test_shift_three_particles_superpose():
    n = 40
    dx = 1.0 / n
    pi = np.array([0.5, 0.5])
    pj = pi + np.array([1.3 * dx, 0.4 * dx])
    pk = pi + np.array([-0.6 * dx, -1.9 * dx])
    sh, vol, dx, h = small_state([pi, pj, pk], n)
    exp_i = shift_pair_term(pi, pj, vol[1], h, dx) + shift_pair_term(pi, pk, vol[2], h, dx)
    np.testing.assert_allclose(sh[0], exp_i, rtol=1e-13)
    exp_j = shift_pair_term(pj, pi, vol[0], h, dx) + shift_pair_term(pj, pk, vol[2], h, dx)
    np.testing.assert_allclose(sh[1], exp_j, rtol=1e-13)


def test_shift_zero_on_the_perfect_lattice():
    n = 40
    o = po.Oracle(nx=n, ny=n, integrator=2)
    o.init_tgv()
    run_eval(o)
    o.shift_sums()
    dx = 1.0 / n
    h = 1.2 * dx
    one_term = 0.5 * h * h * 1.2 * abs(dwdr_analytic(dx, h)) * dx * dx
    typ = o.i32("type")
    sh = np.hypot(o.f64("shift_x"), o.f64("shift_y"))[typ == 2]
    assert np.max(sh) < 1e-13 * one_term


def test_shift_clamp_density_correction_and_eos():
    """Shift_particle_position.f90:108-136 on an interior state with the shift
    coefficient doubled (beta = 1), so that about half the particles hit the
    0.25 dx clamp."""
    n = 32
    o0 = po.Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    rng = np.random.default_rng(4)
    dx = 1.0 / n
    x = o0.f64("g_x")[:N] + rng.uniform(-0.3, 0.3, N) * dx
    y = o0.f64("g_y")[:N] + rng.uniform(-0.3, 0.3, N) * dx
    st = (x, y, o0.f64("g_u")[:N], o0.f64("g_v")[:N], o0.f64("g_rho")[:N] * (1 + rng.normal(0, 2e-3, N)),
          o0.f64("g_mass")[:N])
    o0.close()
    o = po.Oracle(nx=n, ny=n, integrator=2, shift_beta=1.0)
    o.set_fluid_state(*st)
    run_eval(o)
    o.shift_sums()
    sx0, sy0 = o.f64("shift_x"), o.f64("shift_y")
    x0, y0, r0 = o.f64("x"), o.f64("y"), o.f64("rho")
    gx, gy = o.f64("grad_rho_x"), o.f64("grad_rho_y")
    typ, near = o.i32("type"), o.i32("near_surface")
    o.shift()
    x1, y1, r1, p1 = o.f64("x"), o.f64("y"), o.f64("rho"), o.f64("p")
    move = (typ == 2) & (near == 0)
    assert move.sum() == N                      # periodic lattice: no surface at all
    mx = 0.25 * dx
    a = np.sqrt(sx0 * sx0 + sy0 * sy0)
    clamp = move & (a > mx)
    assert 50 < clamp.sum() < move.sum() - 50
    with np.errstate(invalid="ignore", divide="ignore"):
        ex = np.where(clamp, sx0 / a * mx, sx0)
        ey = np.where(clamp, sy0 / a * mx, sy0)
    np.testing.assert_array_equal(x1[move], (x0 + ex)[move])
    np.testing.assert_array_equal(y1[move], (y0 + ey)[move])
    rr = (r0 + gx * ex) + gy * ey
    np.testing.assert_array_equal(r1[move], rr[move])
    dmax = np.sqrt((x1 - x0) ** 2 + (y1 - y0) ** 2)[move]
    assert np.max(dmax) <= mx * (1 + 1e-12)
    pe, _ = eos(r1[move])
    np.testing.assert_array_equal(p1[move], pe)
    ghosts = typ != 2
    np.testing.assert_array_equal(x1[ghosts], x0[ghosts])


def eos(rho):
    import ctypes
    o = po.Oracle()
    p, c = np.zeros(len(rho)), np.zeros(len(rho))
    pp, cc = ctypes.c_double(), ctypes.c_double()
    for k, r in enumerate(rho):
        po.lib().orc_status_equation(o._c, float(r), 0.0, 2, ctypes.byref(pp), ctypes.byref(cc))
        p[k], c[k] = pp.value, cc.value
    o.close()
    return p, c


# --- predictor-corrector -------------------------------------------------------------
@pytest.mark.parametrize("nsteps", [1, 7])
def test_pc_lone_particle_under_gravity_closed_form(nsteps):
    """time_iteration.f90:541-657 with a = (0, -g) and no neighbours: the corrector
    gives v_{n+1} = v_n - g dt and y_{n+1} = y_n + dt (v_n - g dt) - g dt^2 / 2 (the
    trapezoid of the predicted and corrected velocities plus the reference's extra
    dt^2 a / 2), so y_N = y_0 + v_0 N dt - g dt^2 (N^2 + 2N) / 2, x_N = x_0 + u_0 N dt."""
    g, dt = 9.81, 1e-3
    x0, y0, u0, v0 = 0.5, 0.5, 0.3, 0.2
    o = po.Oracle(nx=20, ny=20, integrator=2, dt=dt, gravity=g, mls_on=0)
    o.set_fluid_state(np.array([x0]), np.array([y0]), np.array([u0]), np.array([v0]), np.array([1000.0]),
                      np.array([2.5]))
    for it in range(1, nsteps + 1):
        o.time_step(it)
    N = nsteps
    assert o.f64("g_u")[0] == pytest.approx(u0, abs=1e-15)
    assert o.f64("g_v")[0] == pytest.approx(v0 - N * g * dt, rel=1e-14)
    assert o.f64("g_x")[0] == pytest.approx(x0 + u0 * N * dt, rel=1e-14)
    assert o.f64("g_y")[0] == pytest.approx(y0 + v0 * N * dt - g * dt * dt * (N * N + 2 * N) / 2, rel=1e-14)
    assert o.f64("g_rho")[0] == 1000.0


# --- single-rank local order -----------------------------------------------------------
def local_order_numpy(x, y, typ, h, box_x, box_y):
    """Devide_block_and_send_tasks.f90 on one rank: fluid with y < 0 become ill
    (:75-83); 4h cells from the origin -5 cells, x_n = floor((x - x0)/dx_c + 1),
    row-major cell id, outside the grid ill (:123-162); local order = columns
    ascending, rows ascending, gids ascending (:488-633)."""
    t = typ.copy()
    t[(t == 2) & (y < 0.0)] = 120
    d = (3 + 1) * h
    ox = oy = -(5 * d)
    nx = int(box_x / d + 2 * 5)
    ny = int(box_y / d + 2 * 5)
    xn = np.floor((x - ox) / d + 1).astype(np.int64)
    yn = np.floor((y - oy) / d + 1).astype(np.int64)
    k = (yn - 1) * nx + xn
    ok = (t != 120) & (t != 0)
    inside = (k >= 1) & (k <= nx * ny) & (xn >= 1) & (xn <= nx)
    t[ok & ~inside] = 120
    ok &= inside
    gid = np.flatnonzero(ok)
    order = np.lexsort((gid, yn[gid], xn[gid]))
    return gid[order], t


def test_local_order_numpy_restatement_with_ill_particles():
    n = 40
    o0 = po.Oracle(nx=n, ny=n)
    o0.init_tgv()
    N = n * n
    rng = np.random.default_rng(12)
    dx = 1.0 / n
    x = o0.f64("g_x")[:N] + rng.uniform(-0.49, 0.49, N) * dx
    y = o0.f64("g_y")[:N] + rng.uniform(-0.49, 0.49, N) * dx
    st = [x, y, o0.f64("g_u")[:N], o0.f64("g_v")[:N], o0.f64("g_rho")[:N], o0.f64("g_mass")[:N]]
    o0.close()
    x[3], y[3] = 0.31, -0.004                 # below y = 0: typed 120
    x[17], y[17] = 9.0, 0.5                   # right of the background grid: typed 120
    x[29], y[29] = 0.5, -7.0                  # below the grid (and y < 0)
    o = po.Oracle(nx=n, ny=n, integrator=2)
    o.set_fluid_state(*st)
    o.devide_block(1)                          # no periodic wrap: the ill particles stay where they are
    want, t = local_order_numpy(x, y, np.full(N, 2, dtype=np.int32), 1.2 * dx, 1.0, 1.0)
    assert np.array_equal(o.i32("name"), want)
    assert np.array_equal(np.flatnonzero(o.i32("g_type") == 120), [3, 17, 29])
    assert np.array_equal(np.flatnonzero(t == 120), [3, 17, 29])
    # with the periodic ghosts of a wrapped state the rule is the same over all gids
    o2 = po.Oracle(nx=n, ny=n, integrator=2)
    st2 = [a.copy() for a in st]
    st2[1][[3, 29]] = 0.2
    st2[0][17] = 0.7
    o2.set_fluid_state(*st2)
    o2.period_boundary()
    o2.devide_block(1)
    G = o2.n_global
    want2, _ = local_order_numpy(o2.f64("g_x")[:G], o2.f64("g_y")[:G], o2.i32("g_type")[:G], 1.2 * dx, 1.0, 1.0)
    assert np.array_equal(o2.i32("name"), want2)
    assert G > N


def test_local_order_hand_worked():
    """10x10 box (dx = 0.1, h = 0.12, 4h cells of 0.48 from -2.4): x = 0.05, 0.45
    fall in cell column 6, x = 0.5 in column 7; y = 0.05, 0.1 in row 6, y = 0.9 in
    row 7. Column 6 row 6 holds gids 2, 3; column 6 row 7 gid 1; column 7 gid 0."""
    pts = np.array([(0.5, 0.1), (0.05, 0.9), (0.45, 0.1), (0.05, 0.05)])
    o = po.Oracle(nx=10, ny=10, integrator=2)
    k = len(pts)
    o.set_fluid_state(pts[:, 0], pts[:, 1], np.zeros(k), np.zeros(k), np.full(k, 1000.0), np.full(k, 10.0))
    o.devide_block(1)
    assert list(o.i32("name")) == [2, 3, 1, 0]
