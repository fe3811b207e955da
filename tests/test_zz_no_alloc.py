"""lv_point_values_noalloc: the allocation-free variant of lv_point_values, called through the C ABI.

Its queries and values are staged in the library's static per-device array instead of device
buffers allocated per call.  It must return exactly what lv_point_values returns, on consecutive
calls that reuse the array (the second one smaller, so stale values of the first stay behind it),
and refuse more than LV_NOALLOC_MAX_QUERIES queries without touching the GPU.
"""
import ctypes

import numpy as np
import pytest

from pairs import make_pair

pytestmark = pytest.mark.gpu


def _call(lv, fn, grid, q, fid, comp):
    m = q.size // 2
    out = np.full(m, -7.0)
    rc = getattr(lv.lib(), fn)(grid.ctx, m, q.ctypes.data_as(ctypes.c_void_p), fid, comp,
                               out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


def _queries(rng, m, lo=0.0, hi=1.0):
    q = rng.uniform(lo, hi, size=(m, 2))
    return np.ascontiguousarray(np.concatenate([q[:, 0], q[:, 1]]))


@pytest.mark.parametrize("kind,ic", [("periodic", "tg"), ("walls", "cavity")])
def test_noalloc_point_values_match_the_allocating_entry_point(lv, oracle, gpu, kind, ic):
    print("noalloc point values %s" % kind, flush=True)
    grid, op, x, dr = make_pair(lv, oracle, kind, K=48, seed=0x5EED0900, ic=ic)
    rng = np.random.default_rng(11)
    fid_P, fid_v = lv.FIELDS["P"][0], lv.FIELDS["v"][0]
    # first call: the largest allowed batch; second: a smaller one on another field component
    for m, fid, comp in ((lv.NOALLOC_MAX_QUERIES, fid_P, 0), (1000, fid_v, 1), (1, fid_P, 0)):
        q = _queries(rng, m)
        rc_a, ref = _call(lv, "lv_point_values", grid, q, fid, comp)
        rc_b, got = _call(lv, "lv_point_values_noalloc", grid, q, fid, comp)
        assert rc_a == lv.LV_OK and rc_b == lv.LV_OK, (rc_a, rc_b, lv.lib().lv_last_error(grid.ctx))
        assert np.array_equal(got, ref), m
        assert np.all(np.isfinite(got))
    # the host wrapper gives the same values
    q = _queries(rng, 300)
    pts = np.stack([q[:300], q[300:]], axis=1)
    assert np.array_equal(lv.point_value_noalloc(grid, pts, "P"), lv.point_value(grid, pts, "P"))


def test_noalloc_point_values_refuses_oversized_batches(lv, oracle, gpu):
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=16, seed=0x5EED0901, ic="tg")
    q = _queries(np.random.default_rng(12), lv.NOALLOC_MAX_QUERIES + 1)
    rc, out = _call(lv, "lv_point_values_noalloc", grid, q, lv.FIELDS["P"][0], 0)
    assert rc == lv.LV_E_INVALID
    assert b"LV_NOALLOC_MAX_QUERIES" in lv.lib().lv_last_error(grid.ctx)
    assert np.all(out == -7.0)
    # an unknown field component and zero queries
    rc, _ = _call(lv, "lv_point_values_noalloc", grid, q[:20], lv.FIELDS["P"][0], 3)
    assert rc == lv.LV_E_INVALID
    rc, _ = _call(lv, "lv_point_values_noalloc", grid, np.zeros(0), lv.FIELDS["P"][0], 0)
    assert rc == lv.LV_OK
