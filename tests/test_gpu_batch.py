"""The batched form (sph2d_batch): independent contexts advanced together, the
repeating and MLS steps as hipGraph launches (one launch per kernel over all
instances for the predictor-corrector, one branch per instance for the
RK-variant). Each instance must come out bit-identical to the same instance run
alone, after graph replays separated by eager steps (step 1; the MLS step 20 for
the RK-variant) and by host synchronisations."""
import sys

import numpy as np
import pytest

import sph2d
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(32, 100.0), (48, 400.0), (40, 60.0)]


def make(precision, integrator):
    sims = []
    for n, re in CASES:
        prm = sph2d.tgv_params(n, precision=precision, integrator=integrator, nu=1.0 / re, sample_every=100)
        sim = sph2d.Simulation(prm)
        sim.init_tgv()
        sims.append(sim)
    return sims


@pytest.mark.parametrize("precision,integrator", [(sph2d.FP64, sph2d.INTEGRATOR_PC),
                                                  (sph2d.FP32_MIXED, sph2d.INTEGRATOR_PC),
                                                  (sph2d.FP64, sph2d.INTEGRATOR_RK)])
def test_batch_is_bit_identical_to_each_instance_alone(precision, integrator):
    say(f"test_batch_is_bit_identical_to_each_instance_alone precision={precision} integrator={integrator}")
    bat = make(precision, integrator)
    ref = make(precision, integrator)
    batch = sph2d.Batch(bat)
    done = 0
    for chunk in (10, 11, 4):                    # checkpoints after steps 10, 21 and 25
        batch.run(done + 1, chunk)
        for s in ref:
            s.run(done + 1, chunk)
        done += chunk
        for k, (a, b) in enumerate(zip(bat, ref)):
            ga, gb = a.download_state(), b.download_state()
            for f in ("x", "y", "u", "v", "rho", "p"):
                assert np.array_equal(ga[f], gb[f]), (done, k, f)
        say(f"step {done}: identical; graph nodes/replays {batch.info()}")
    nodes, replays = batch.info()
    # steps 2..25; the MLS step 20 is replayed too when the launches are fused (PC)
    assert nodes >= 15 and replays == (24 if integrator == sph2d.INTEGRATOR_PC else 23)
    ke = [s.sample()[0] for s in bat]
    assert all(0.9 < k < 1.0 for k in ke)
    batch.close()
    for s in bat + ref:
        s.close()


def test_batch_rejects_mismatched_schedules():
    say("test_batch_rejects_mismatched_schedules")
    a = sph2d.Simulation(sph2d.tgv_params(32, mls_every=20))
    b = sph2d.Simulation(sph2d.tgv_params(32, mls_every=10))
    with pytest.raises(sph2d.SphError):
        sph2d.Batch([a, b])
    with pytest.raises(sph2d.SphError):
        sph2d.Batch([a, a])
    a.close()
    b.close()


def test_autotuned_batch_stays_bit_identical():
    """The autotuner re-lays the fused launches 16 times (instance order, block
    order, tile shape of the mixed neighbour passes) while it advances the
    instances; every configuration must leave each instance bit-identical to
    running it alone."""
    say("test_autotuned_batch_stays_bit_identical")
    bat = make(sph2d.FP32_MIXED, sph2d.INTEGRATOR_PC)
    ref = make(sph2d.FP32_MIXED, sph2d.INTEGRATOR_PC)
    batch = sph2d.Batch(bat)
    batch.run(1, 5)
    ms, chosen = batch.autotune(6, 2)                # steps 6..53
    say(f"autotune ms/step {ms}, chosen {sph2d.Batch.CONFIGS[chosen]}")
    assert np.all(ms > 0) and len(ms) == 16 and 0 <= chosen < 16 and batch.config == chosen
    batch.run(54, 6)
    for s in ref:
        s.run(1, 59)
    for k, (a, b) in enumerate(zip(bat, ref)):
        ga, gb = a.download_state(), b.download_state()
        for f in ("x", "y", "u", "v", "rho", "p"):
            assert np.array_equal(ga[f], gb[f]), (k, f)
    batch.close()
    for s in bat + ref:
        s.close()


def test_batch_of_walled_tanks_uses_branches_and_stays_identical():
    """Contexts with fixed walls and gravity cannot share fused launches; the batch
    replays one graph with each instance's step as a branch, bit-identical to
    running each alone."""
    say("test_batch_of_walled_tanks_uses_branches_and_stays_identical")

    def tanks():
        out = []
        for n in (30, 36):
            prm = sph2d.tgv_params(n, walls=4, periodic=0, gravity=9.81, sample_every=100)
            sim = sph2d.Simulation(prm)
            sim.init_tgv()
            out.append(sim)
        return out
    bat, ref = tanks(), tanks()
    batch = sph2d.Batch(bat)
    batch.run(1, 12)
    for s in ref:
        s.run(1, 12)
    nodes, replays = batch.info()
    assert replays == 11 and nodes > 30
    for a, b in zip(bat, ref):
        ga, gb = a.download_state(), b.download_state()
        for f in ("x", "y", "u", "v", "rho", "p"):
            assert np.array_equal(ga[f], gb[f]), f
    batch.close()
    for s in bat + ref:
        s.close()


def test_uneven_64_instance_batch():
    """The 64-instance workload of uneven sizes (sides 64..512 log-uniform, Re
    50..1000, ~3.5M particles; sph2d.uneven_tgv_instances(64, seed=3)), FP32-mixed
    predictor-corrector, autotuned, 500 steps:

    * the smallest instance (64^2, Re 901), started from the oracle's initial state,
      against the oracle (FP64) after 10 steps: positions within 2e-8, u and v within
      2e-8 and rho within 5e-9 of the field's largest magnitude (measured 6.1e-9,
      5.0e-9, 1.1e-9) for all but 1% of the particles, and every particle within
      5e-6 (positions) and 1e-6 (u, v, rho). Pairs of the moving lattice cross the
      3h cutoff every step (lattice pairs at sqrt(13) dx = 3.606 dx start 0.006 dx
      outside it); one that crosses within the runs' ~5e-9 difference of a step
      boundary joins one run's neighbour set a step before the other's, and since
      the truncated kernel's gradient at 3h is not zero, the shift of the particles
      near that pair differs by ~1e-6 from then on. The runs part there and only
      there: every other particle keeps the tight bars;
    * every tile shape and launch layout advances every instance identically: after
      the autotuner's 16 trials the largest and the smallest instances equal the
      same instances run alone;
    * at step 500 every instance's state is finite and its kinetic energy lies within
      3% of the analytic decay exp(-16 pi^2 nu t) (the weakly compressible field
      starts with a short acoustic transient that the analytic solution lacks)."""
    say("test_uneven_64_instance_batch")
    sides, res = sph2d.uneven_tgv_instances(64, seed=3)
    assert len(sides) == 64 and min(sides) >= 64 and max(sides) <= 512
    k_min, k_max = int(np.argmin(sides)), int(np.argmax(sides))

    def instance(k):
        prm = sph2d.tgv_params(sides[k], integrator=sph2d.INTEGRATOR_PC, precision=sph2d.FP32_MIXED, nu=1.0 / res[k])
        return sph2d.Simulation(prm)

    sims = [instance(k) for k in range(64)]
    for s in sims:
        s.init_tgv()
    n = sides[k_min]
    o = Oracle(nx=n, ny=n, integrator=2, dt=sims[k_min].params.dt, nu=1.0 / res[k_min])
    o.init_tgv()
    st = tuple(o.f64(f)[:n * n].copy() for f in ("g_x", "g_y", "g_u", "g_v", "g_rho", "g_mass"))
    sims[k_min].upload_state(*st)
    alone = {k: instance(k) for k in (k_min, k_max)}
    alone[k_min].upload_state(*st)
    alone[k_max].init_tgv()
    say(f"64 instances, {sum(q * q for q in sides)} particles, sides {min(sides)}..{max(sides)}")
    batch = sph2d.Batch(sims)
    batch.run(1, 10)
    for it in range(1, 11):
        o.time_step(it)
    g = sims[k_min].download_state()
    errs, tail = {}, {}
    for f, of in (("x", "g_x"), ("y", "g_y"), ("u", "g_u"), ("v", "g_v"), ("rho", "g_rho")):
        r = o.f64(of)[:n * n]
        e = np.abs(g[f] - r) / (1.0 if f in ("x", "y") else float(np.max(np.abs(r))))
        errs[f] = float(np.max(e))
        tail[f] = float(np.quantile(e, 0.99))
    say(f"smallest instance ({n}^2) vs the oracle after 10 steps: max {errs}, 99th percentile {tail}")
    for f, bar, loose in (("x", 2e-8, 5e-6), ("y", 2e-8, 5e-6), ("u", 2e-8, 1e-6), ("v", 2e-8, 1e-6),
                          ("rho", 5e-9, 1e-6)):
        assert tail[f] < bar, (f, tail[f])
        assert errs[f] < loose, (f, errs[f])
    o.close()
    ms, chosen = batch.autotune(11, 5)                   # steps 11..106
    say(f"autotune ms/step {np.round(ms, 3).tolist()}, chosen {sph2d.Batch.CONFIGS[chosen]}")
    assert np.all(ms > 0)
    for k, a in alone.items():
        a.run(1, 106)
        ga, gb = sims[k].download_state(), a.download_state()
        for f in ("x", "y", "u", "v", "rho"):
            assert np.array_equal(ga[f], gb[f]), (k, f)
    batch.run(107, 394)
    worst = 0.0
    for k, s in enumerate(sims):
        smp = s.sample()
        t = 500 * s.params.dt
        ka = np.exp(-16 * np.pi ** 2 * s.params.nu * t)
        assert np.all(np.isfinite(smp)), k
        worst = max(worst, abs(smp[0] - ka) / ka)
        assert abs(smp[0] - ka) / ka < 0.03, (k, sides[k], res[k], smp[0], ka)
    say(f"step 500: largest |KE/KE0 - analytic| / analytic over the 64 instances {worst:.3e}")
    batch.close()
    for s in sims + list(alone.values()):
        s.close()
