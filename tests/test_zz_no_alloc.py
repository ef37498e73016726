"""kazan_cost_model_eval_noalloc (include/kazan/kazan_tools.h): the cost
model's inference staged through the library's static device array instead
of device allocations. Its predictions must equal kazan_cost_model_eval's
bit for bit (the same kernel on the same inputs) on two calls in a row, the
second with fewer rows, so that the first call's contents left in the array
cannot leak into the second's results; rows beyond the array's limit are
refused before anything runs."""
import ctypes

import numpy as np
import pytest

from kazan_amd.balancer import model as M

MAX_ROWS = 245760


def _fn(driver, name):
    fn = getattr(driver.lib, name)
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    return fn


def _blob(seed):
    r = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[M.P_B1:M.P_W4] = r.normal(0, 0.1, M.P_W4 - M.P_B1)
    p[M.P_B4] = 4.0
    return M.pack(p, r.normal(0, 1, 16), r.uniform(0.5, 2, 16), 0.5)


def _run(fn, blob, feat):
    feat = np.ascontiguousarray(feat, np.float32)
    out = np.full(len(feat), np.nan, np.float32)
    rc = fn(blob, len(blob), feat.ctypes.data, len(feat), out.ctypes.data)
    assert rc == 0, rc
    return out


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point(driver):
    plain = _fn(driver, "kazan_cost_model_eval")
    noalloc = _fn(driver, "kazan_cost_model_eval_noalloc")
    rng = np.random.default_rng(11)
    blob1, blob2 = _blob(5), _blob(6)
    for blob, rows in ((blob1, 20000), (blob2, 1000)):
        print("noalloc eval: %d rows" % rows, flush=True)
        feat = rng.normal(0, 2, (rows, 16)).astype(np.float32)
        want = _run(plain, blob, feat)
        got = _run(noalloc, blob, feat)
        assert np.isfinite(got).all()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # the largest size it accepts
    feat = rng.normal(0, 2, (MAX_ROWS, 16)).astype(np.float32)
    print("noalloc eval: %d rows (the limit)" % MAX_ROWS, flush=True)
    np.testing.assert_array_equal(_run(noalloc, blob1, feat).view(np.uint32), _run(plain, blob1, feat).view(np.uint32))


def test_noalloc_refuses_rows_beyond_its_array(driver):
    noalloc = _fn(driver, "kazan_cost_model_eval_noalloc")
    blob = _blob(5)
    out = np.zeros(1, np.float32)
    feat = np.zeros((1, 16), np.float32)
    # the row count is checked before any pointer is read
    assert noalloc(blob, len(blob), feat.ctypes.data, MAX_ROWS + 1, out.ctypes.data) == -5
    assert noalloc(blob[:-4], len(blob) - 4, feat.ctypes.data, 1, out.ctypes.data) == -1
