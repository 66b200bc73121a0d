"""The allocation-free endpoint dedup (bt_read_endpoints_noalloc) through the
C ABI against bt_read_endpoints, which allocates its table per call: equal
index lists and counts on two calls in a row, the second on a smaller graph
(the shared device scratch keeps the first call's contents, so anything not
cleared would show), after iterations, with a short output buffer, and the
refusal of a graph above the scratch's path limit."""
import ctypes

import numpy as np
import pytest

from bundling_thingy_amd import api, workloads

pytestmark = pytest.mark.gpu

_U32P = ctypes.POINTER(ctypes.c_uint32)


def _call(fn, gpu, capacity):
    out = np.zeros(max(capacity, 1), dtype=np.uint32)
    count = ctypes.c_uint32(0xdeadbeef)
    rc = fn(gpu._h, out.ctypes.data_as(_U32P) if capacity else None, capacity, ctypes.byref(count))
    return rc, out[:min(capacity, count.value)], count.value


def _both(gpu):
    L = api.lib()
    rc0, _, n0 = _call(L.bt_read_endpoints, gpu, 0)
    rc1, _, n1 = _call(L.bt_read_endpoints_noalloc, gpu, 0)
    assert rc0 == rc1 == api.BT_E_CAPACITY and n0 == n1 > 0
    rc0, want, _ = _call(L.bt_read_endpoints, gpu, n0)
    rc1, got, n = _call(L.bt_read_endpoints_noalloc, gpu, n0)
    assert rc0 == rc1 == 0 and n == n0
    np.testing.assert_array_equal(got, want)
    return want


def test_noalloc_endpoints_equal_allocating_twice_smaller_second(bt, orc):
    big = workloads.plain_gpu_graph()
    small = workloads.plain_cpu_graph()
    opts = bt.Options(spacing=0.002, radius=0.01, sigma=2.0, density=1024)
    with bt.Gpu(opts) as gpu:
        gpu.upload(big)
        want = _both(gpu)
        np.testing.assert_array_equal(want, orc.endpoints(gpu.points()[0], gpu.paths()))
        gpu.iterate(2)
        np.testing.assert_array_equal(_both(gpu), want)
    with bt.Gpu(workloads.reference_cpu_options()) as gpu:
        gpu.upload(small)
        want_small = _both(gpu)
        assert len(want_small) < len(want)
        np.testing.assert_array_equal(want_small,
                                      orc.endpoints(gpu.points()[0], gpu.paths()))
        # a short buffer: the smallest indices, and BT_E_CAPACITY
        rc, got, n = _call(api.lib().bt_read_endpoints_noalloc, gpu, 5)
        assert rc == api.BT_E_CAPACITY and n == len(want_small)
        np.testing.assert_array_equal(got, want_small[:5])


def test_noalloc_endpoints_refuses_above_path_limit(bt):
    k = (1 << 18) + 1  # one path above BT_NOALLOC_MAX_PATHS
    rng = np.random.default_rng(9)
    a = rng.uniform(0.0, 1.0, (k, 2)).astype(np.float32)
    pts = np.empty((2 * k, 2), np.float32)
    pts[0::2] = a
    pts[1::2] = a + np.float32(0.003)
    g = bt.Graph(pts, np.arange(0, 2 * k + 1, 2, dtype=np.uint32))
    with bt.Gpu(bt.Options(map_size=201, neighbor_radius=4)) as gpu:
        gpu.upload(g)
        rc, _, n = _call(api.lib().bt_read_endpoints_noalloc, gpu, 16)
        assert rc == api.BT_E_CAPACITY and n == 0
        assert b"at most 262144" in api.lib().bt_last_error(gpu._h)
        assert len(gpu.endpoints()) > 0  # the allocating form takes it
