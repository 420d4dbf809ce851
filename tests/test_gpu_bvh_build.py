"""The GPU binned-SAH build (bvh_build.hip, pg_bvh_build_gpu) against the reference's
topology: node arrays and triangle order bit-identical to the host build (itself pinned
to the reference's BVHNode::build, Geometry.h:274-446), on scenes with failed splits
(clusters of equal centroids), signed-zero ties in the bounds (the reference's own
build, tests/golden/ref_bvh_signed_zero.npz), tiny triangle counts, and the
~1.09M-triangle interior (the reference's topology hash). A scene uploaded with the
GPU builder renders the same film as with the host builder."""
import ctypes as C
import hashlib
import json

import numpy as np
import pytest

from conftest import GOLDEN, clustered_scene

pytestmark = pytest.mark.gpu


def _build(pg, device, pos, host=False):
    lib = device.lib
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    nn = C.c_uint32()
    nodes = np.empty((max(2 * n - 1, 1), 8), np.float32)
    idx = np.empty(n, np.int32)
    ms = C.c_double()
    fp, ip = pos.ctypes.data_as(C.POINTER(C.c_float)), idx.ctypes.data_as(C.POINTER(C.c_int32))
    if host:
        rc = lib.pg_bvh_build_host(fp, n, C.byref(nn), nodes.ctypes.data_as(C.POINTER(C.c_float)), ip)
    else:
        rc = lib.pg_bvh_build_gpu(device.ctx, fp, n, C.byref(nn), nodes.ctypes.data_as(C.POINTER(C.c_float)), ip,
                                  C.byref(ms))
    assert rc == 0, lib.pg_last_error(device.ctx)
    return nodes[: nn.value], idx, ms.value


@pytest.mark.parametrize("which", ["cornell", "interior100", "window101", "clustered"])
def test_gpu_build_equals_host_build(pg, device, which):
    data = {"cornell": lambda: pg.generate_scene(1, 0), "interior100": lambda: pg.generate_scene(100, 0),
            "window101": lambda: pg.generate_scene(101, 3), "clustered": lambda: clustered_scene(pg)}[which]()
    g_nodes, g_idx, _ = _build(pg, device, data.positions)
    h_nodes, h_idx, _ = _build(pg, device, data.positions, host=True)
    assert g_nodes.shape == h_nodes.shape, (g_nodes.shape, h_nodes.shape)
    assert np.array_equal(g_nodes.view(np.uint32), h_nodes.view(np.uint32))
    assert np.array_equal(g_idx, h_idx)


def test_gpu_build_signed_zero_ties_match_reference(pg, device):
    g = np.load(GOLDEN / "ref_bvh_signed_zero.npz")
    nodes, idx, _ = _build(pg, device, g["positions"])
    assert np.array_equal(nodes.view(np.uint32), g["nodes"].view(np.uint32))
    assert np.array_equal(idx, g["tri_index"])


@pytest.mark.parametrize("n", [1, 2, 4, 5, 6, 9, 33])
def test_gpu_build_small_and_degenerate(pg, device, n):
    rng = np.random.default_rng(n)
    pos = rng.uniform(-1, 1, (n, 3, 3)).astype(np.float32)
    for p in (pos, np.repeat(pos[:1], n, 0)):  # random, then n copies of one triangle (no split possible)
        g_nodes, g_idx, _ = _build(pg, device, p)
        h_nodes, h_idx, _ = _build(pg, device, p, host=True)
        assert np.array_equal(g_nodes.view(np.uint32), h_nodes.view(np.uint32)) and np.array_equal(g_idx, h_idx)


def test_gpu_build_full_interior_matches_reference_hash(pg, device):
    ref = json.load(open(GOLDEN / "ref_topology_config2.json"))
    data = pg.generate_scene(2, 0)
    nodes, idx, ms = _build(pg, device, data.positions)
    assert nodes.shape[0] == ref["n_nodes"]
    assert hashlib.sha256(nodes.tobytes() + idx.tobytes()).hexdigest() == ref["topology_sha256"]
    times = [_build(pg, device, data.positions)[2] for _ in range(3)]
    print(f"GPU BVH build of {data.n_tris} triangles: {min(times):.2f} ms (first {ms:.2f} ms)", flush=True)


def test_scene_upload_with_gpu_builder_renders_the_same(pg, device, small_interior):
    films = []
    for builder in (0, 1):
        assert device.lib.pg_set_bvh_builder(device.ctx, builder) == 0
        sc = pg.Scene(device, small_interior)
        rt = pg.RayTracer(sc, seed=3, max_depth=12)
        rt.clear()
        rt.render(2)
        films.append(rt.film()[0])
        info = sc.bvh_info()
        assert info["nodes"] > 1
    device.lib.pg_set_bvh_builder(device.ctx, 0)
    assert np.array_equal(films[0].view(np.uint32), films[1].view(np.uint32))
    assert device.lib.pg_set_bvh_builder(device.ctx, 7) == pg._lib.PG_ERR_INVALID
