"""sph2d_reorder_particle_index_noalloc (the no-allocation variant of
sph2d_reorder_particle_index, scratch in the library's static per-GPU array)
against the allocating entry point through the C-ABI: permutation, counts and
every permuted column bit-exact, on calls in a row whose second input is
smaller (the array keeps the last call's contents), with and without a caller
perm array, and the size limit rejected before any device access."""
import ctypes

import numpy as np
import pytest

import sph2d
from test_gpu_reorder import case

pytestmark = pytest.mark.gpu


def same(a, b):
    assert np.array_equal(a[6], b[6]), (a[6], b[6])
    if a[5] is not None and b[5] is not None:
        assert np.array_equal(a[5], b[5])
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u, v)
    for u, v in zip(a[3] + a[4], b[3] + b[4]):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("with_perm", [True, False])
def test_noalloc_equals_the_allocating_entry_point_on_calls_in_a_row(with_perm):
    print(f"test_noalloc_equals_the_allocating_entry_point_on_calls_in_a_row with_perm={with_perm}", flush=True)
    for n, nf, seed in ((1_500_000, 1_200_000, 11), (70_001, 50_000, 12), (257, 0, 13)):
        nf, x, y, t, fc, ic = case(n, nf, seed)
        ref = sph2d.reorder_particle_index(nf, 1.0, 1.0, x, y, t, fc, ic)
        got = sph2d.reorder_particle_index(nf, 1.0, 1.0, x, y, t, fc, ic, no_alloc=True, with_perm=with_perm)
        same(got, ref)
        assert int(got[6][5]) > 0 or n < 300


def test_noalloc_rejects_sizes_beyond_its_scratch():
    print("test_noalloc_rejects_sizes_beyond_its_scratch", flush=True)
    x = sph2d.DeviceArray(np.zeros(4))
    y = sph2d.DeviceArray(np.zeros(4))
    t = sph2d.DeviceArray(np.full(4, 2, dtype=np.int32))
    counts = np.zeros(6, dtype=np.int32)
    rc = sph2d.lib().sph2d_reorder_particle_index_noalloc((1 << 21) + 1, 0, 1.0, 1.0, x.ptr, y.ptr, t.ptr, 0,
                                                          None, 0, None, None,
                                                          counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                          None)
    assert rc == sph2d.ERR_CAPACITY
    assert b"2097152" in sph2d.lib().sph2d_last_error()
    for d in (x, y, t):
        d.free()
