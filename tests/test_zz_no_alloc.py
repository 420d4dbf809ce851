"""pg_trace_closest_noalloc, the entry point that allocates nothing (rays and hits stage
through one fixed device array per GPU), against pg_trace_closest: bit-identical
closest hits on two calls in a row, the first at the 2^18-ray limit and the second
smaller (the array keeps the first call's contents), and a clear error above the
limit. Called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from conftest import seeded_rays

pytestmark = pytest.mark.gpu

HIT = np.dtype([("id", "u4"), ("t", "f4"), ("alpha", "f4"), ("beta", "f4"), ("gamma", "f4")])


def _call(dev, fn, rays):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    hits = np.zeros(rays.shape[0], HIT)
    code = fn(dev.ctx, rays.ctypes.data_as(C.POINTER(C.c_float)), rays.shape[0], hits.ctypes.data_as(C.c_void_p))
    return code, hits


def test_trace_closest_noalloc_matches_trace_closest(pg, device, small_interior):
    from oracle.pg_oracle import OracleScene
    pg.Scene(device, small_interior)
    lib = device.lib
    rays = seeded_rays(OracleScene(small_interior), small_interior, 1 << 18, seed=31)
    for n in (1 << 18, 10_007):
        r = rays[:n] if n == (1 << 18) else rays[::-1][:n].copy()
        c0, ref = _call(device, lib.pg_trace_closest, r)
        c1, got = _call(device, lib.pg_trace_closest_noalloc, r)
        assert c0 == 0 and c1 == 0, (c0, c1, lib.pg_last_error(device.ctx))
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), n
        assert (got["id"] != 0xFFFFFFFF).mean() > 0.3
    code, _ = _call(device, lib.pg_trace_closest_noalloc, np.zeros(((1 << 18) + 1, 6), np.float32))
    assert code == pg._lib.PG_ERR_INVALID
    assert b"2^18" in lib.pg_last_error(device.ctx)
