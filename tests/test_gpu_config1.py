"""Taylor-Green at the full sizes of the 1M- and 4.2M-particle workloads.

* 1024^2 (1,048,576 fluid + 16,448 ghost particles, ~19M pairs): one FP64
  evaluation against the oracle run on the host, rates per particle within 1e-12
  relative to max(|ref_i|, 1e-3 sum_k |term_ik|); the FP32-mixed evaluation of the
  same state against the same oracle, every output within 1e-4 of its field norm.
* 2048^2 (4,194,304 + 32,832): the FP32-mixed form through size-independent
  properties -- the exact lattice pair count (18 N + G/2), ghost and local counts,
  a finite state after 20 steps, and its rates against the FP64 GPU form of the
  same state (per output, 1e-4 of the field norm).
"""
import ctypes
import sys

import numpy as np
import pytest

import sph2d
from oracle.pyoracle import Oracle
from test_gpu_parity import particle_rel_err, rel_err, say
from test_oracle_physics import lattice_pair_count

pytestmark = pytest.mark.gpu

TOL = 1e-12
MIXED_TOL = 1e-4


@pytest.fixture(scope="module")
def oracle_1024():
    n = 1024
    say(f"oracle {n}^2 evaluation on the host")
    o = Oracle(nx=n, ny=n, integrator=2, dt=1e-4 * 100.0 / n)
    o.init_tgv()
    o.period_boundary()
    o.devide_block(1)
    o.single_step_compute(1)
    yield o
    o.close()


def gpu_eval(n, precision):
    prm = sph2d.tgv_params(n, integrator=2, precision=precision)
    sim = sph2d.Simulation(prm)
    sim.init_tgv()
    sim.distribute_period_boundary()
    sim.devide_block(1)
    sim.single_step_compute(1)
    return sim


def test_1024_fp64_evaluation_per_particle_vs_oracle(oracle_1024):
    say("test_1024_fp64_evaluation_per_particle_vs_oracle")
    o = oracle_1024
    sim = gpu_eval(1024, sph2d.FP64)
    d = sim.diagnostics()
    assert d.n_ghost == 16448 and d.n_local == o.n_local
    assert np.array_equal(sim.local("name"), o.i32("name"))
    cnt = ctypes.c_int64()
    assert sph2d.lib().sph2d_download_pairs(sim._c, 0, None, None, ctypes.byref(cnt)) == 0
    assert cnt.value == o.pair_number == lattice_pair_count(1024)
    errs = {f: particle_rel_err(sim.local(f), o.f64(f), o.f64("mag_" + f)) for f in ("drhodt", "ax", "ay")}
    say("1024^2 FP64 per-particle rate errors:", errs)
    for f, e in errs.items():
        assert e < TOL, (f, e)
    for f in ("fra_v", "grad_rho_x", "grad_rho_y", "tau_xx", "tau_yy"):
        assert rel_err(sim.local(f), o.f64(f)) < TOL, f
    sim.close()


def test_1024_mixed_evaluation_vs_oracle(oracle_1024):
    say("test_1024_mixed_evaluation_vs_oracle")
    o = oracle_1024
    sim = gpu_eval(1024, sph2d.FP32_MIXED)
    assert np.array_equal(sim.local("name"), o.i32("name"))
    groups = {"fra_v": ["fra_v"], "sum_dwdx": ["sum_dwdx_x", "sum_dwdx_y"],
              "tau": ["tau_xx", "tau_xy", "tau_yy"], "grad_rho": ["grad_rho_x", "grad_rho_y"],
              "drhodt": ["drhodt"], "a": ["ax", "ay"]}
    worst = {}
    for key, fs in groups.items():
        ref = {f: o.f64(f) for f in fs}
        scale = max(np.max(np.abs(r)) for r in ref.values())
        for f in fs:
            worst[f] = float(np.max(np.abs(sim.local(f) - ref[f])) / scale)
    say("1024^2 mixed-form errors vs FP64 oracle:", worst)
    for f, e in worst.items():
        assert e < MIXED_TOL, (f, e)
    sim.close()


def test_2048_mixed_properties():
    say("test_2048_mixed_properties")
    n = 2048
    fp = gpu_eval(n, sph2d.FP64)
    mx = gpu_eval(n, sph2d.FP32_MIXED)
    d = mx.diagnostics()
    assert d.n_ghost == 32832 and d.n_local == n * n + 32832
    cnt = ctypes.c_int64()
    assert sph2d.lib().sph2d_download_pairs(mx._c, 0, None, None, ctypes.byref(cnt)) == 0
    assert cnt.value == lattice_pair_count(n)
    assert np.array_equal(mx.local("name"), fp.local("name"))
    worst = {}
    for f in ("drhodt", "ax", "ay", "fra_v", "grad_rho_x", "tau_xx"):
        worst[f] = rel_err(mx.local(f), fp.local(f))
    say("2048^2 mixed vs FP64 GPU:", worst)
    for f, e in worst.items():
        assert e < MIXED_TOL, (f, e)
    fp.close()
    mx.run(1, 20)
    s = mx.sample()
    assert np.all(np.isfinite(s)) and abs(s[0] - 1.0) < 1e-3
    st = mx.download_state()
    assert all(np.all(np.isfinite(st[k])) for k in ("x", "y", "u", "v", "rho"))
    d = mx.diagnostics()
    assert d.error_flags == 0 and d.rebin_needed == 0
    mx.close()
