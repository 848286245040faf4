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
    """The autotuner re-lays the fused launches four times (instance order, block
    order) while it advances the instances; every configuration must leave each
    instance bit-identical to running it alone."""
    say("test_autotuned_batch_stays_bit_identical")
    bat = make(sph2d.FP32_MIXED, sph2d.INTEGRATOR_PC)
    ref = make(sph2d.FP32_MIXED, sph2d.INTEGRATOR_PC)
    batch = sph2d.Batch(bat)
    batch.run(1, 5)
    ms, chosen = batch.autotune(6, 4)                # steps 6..25
    say(f"autotune ms/step {ms}, chosen {sph2d.Batch.CONFIGS[chosen]}")
    assert np.all(ms > 0) and 0 <= chosen < 4
    batch.run(26, 6)
    for s in ref:
        s.run(1, 31)
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
