"""Pins the oracle's Reorder_Particle_Index
(Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/Reorder_Particle_Index.f90:54-287):
a hand-worked case, the ill test at its exact bounds (REAL(4) -1.0E-10 below, the
box edge above), and on random inputs the defining properties of a stable
partition (every kept particle once, categories in the order fluid, wavemaker,
fixed ghost, inlet, outlet, source order kept inside each category)."""
import numpy as np

from oracle import pyoracle as po

KEEP = (2, 130, -3, -201, -202)


def test_hand_worked_case():
    #        0    1    2    3     4    5     6     7    8    9
    t = [2, -3, 130, 2, -101, 2, -202, -201, 120, 2]
    x = [0.5, 0.1, 0.2, 1.5, 0.3, 0.0, 0.4, 0.6, 0.7, 0.9]
    y = [0.5] * 10
    v = [10.0 * i for i in range(10)]
    xs, ys, ts, (vs,), _, perm, counts = po.reorder_particle_index(10, 1.0, 1.0, x, y, t, [v])
    # particle 3 is fluid outside the box -> ill (dropped); 4 (periodic) and 8 (ill) dropped
    assert list(perm) == [0, 5, 9, 2, 1, 7, 6]
    assert list(counts) == [3, 1, 1, 1, 1, 7]
    assert list(ts) == [2, 2, 2, 130, -3, -201, -202]
    assert list(vs) == [0.0, 50.0, 90.0, 20.0, 10.0, 70.0, 60.0]


def test_ill_bounds_are_the_references():
    lo = float(np.float32(-1.0e-10))            # REAL(4) literal widened
    xs = [lo, np.nextafter(lo, 1.0), -1.0e-10, 0.0, 1.0, np.nextafter(1.0, 2.0), 0.5, 0.5]
    ys = [0.5, 0.5, 0.5, 0.5, 0.5, 0.5, lo, 2.0]
    t = [2] * 8
    _, _, _, _, _, perm, counts = po.reorder_particle_index(8, 1.0, 1.0, xs, ys, t)
    assert list(perm) == [1, 2, 3, 4]
    assert counts[0] == 4
    # fluid beyond n_fluid is not tested for the box (the reference loops to Particle_ture_number)
    _, _, _, _, _, perm, _ = po.reorder_particle_index(0, 1.0, 1.0, xs, ys, t)
    assert list(perm) == list(range(8))


def test_random_stable_partition_properties():
    rng = np.random.default_rng(9)
    n = 5000
    types = rng.choice([2, 2, 2, 130, -3, -201, -202, -101, 120, -4], size=n)
    x = rng.uniform(-0.1, 1.1, n)
    y = rng.uniform(-0.1, 1.1, n)
    nf = 4000
    ill = (np.arange(n) < nf) & (types == 2) & ((x <= -1e-10) | (x > 1.0) | (y <= -1e-10) | (y > 1.0))
    eff = np.where(ill, 120, types)
    xs, ys, ts, _, (ids,), perm, counts = po.reorder_particle_index(nf, 1.0, 1.0, x, y, types,
                                                                     i32_cols=[np.arange(n)])
    assert np.array_equal(ids, perm)
    expect = np.concatenate([np.nonzero(eff == c)[0] for c in KEEP])
    assert np.array_equal(perm, expect)
    assert list(counts[:5]) == [int(np.sum(eff == c)) for c in KEEP]
    assert np.array_equal(xs, x[perm]) and np.array_equal(ts, eff[perm])
