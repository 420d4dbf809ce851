"""Batches of views (pg_batch_*): every view of a batch equals, bit for bit, the same
view rendered alone with the same seed; hipGraph replays equal eager passes; chunking
and autotuning do not change the images."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(data, w, h, **over):
    c = dict(data.camera)
    c.update(width=w, height=h, **over)
    return c


@pytest.fixture(scope="module")
def views(pg, cornell, small_interior):
    return [
        (0, _cam(cornell, 96, 80)),
        (1, _cam(small_interior, 160, 90)),
        (0, _cam(cornell, 64, 64, **{"from": [0.3, 1.2, 3.0], "fov": 50.0})),
        (1, _cam(small_interior, 48, 120, **{"to": [1.0, 1.5, -3.0]})),
    ]


def _batch(pg, device, cornell, small_interior, views, **cfg):
    b = pg.Batch(device)
    assert b.add_scene(cornell) == 0
    assert b.add_scene(small_interior) == 1
    for sid, cam in views:
        b.add_view(sid, cam)
    b.configure(**cfg)
    b.clear()
    return b


def _single(pg, device, data, cam, spp, seed):
    sc = pg.Scene(device, dataclasses.replace(data, camera=cam))
    rt = pg.RayTracer(sc, seed=seed, max_depth=12)
    rt.clear()
    rt.render(spp)
    return rt.film()[0]


def test_batch_views_equal_single_renders(pg, device, cornell, small_interior, views):
    b = _batch(pg, device, cornell, small_interior, views, use_graph=1)
    b.render(3, seed=7, max_depth=12)
    films = [b.film(v)[0] for v in range(len(views))]
    st = b.stats()
    assert st["views"] == 4 and st["scenes"] == 2 and st["chunks"] == 1 and st["graph_nodes"] > 0
    b.close()
    scenes = [cornell, small_interior]
    for v, (sid, cam) in enumerate(views):
        ref = _single(pg, device, scenes[sid], cam, 3, 7)
        assert np.array_equal(films[v], ref), f"view {v} differs from its single-view render"
        assert films[v].sum() > 0


def test_graph_replay_equals_eager_with_work_between(pg, device, cornell, small_interior, views):
    eager = _batch(pg, device, cornell, small_interior, views, use_graph=0)
    eager.render(4, seed=3, max_depth=12)
    e_f = [eager.film(v)[0] for v in range(len(views))]
    eager.close()
    g = _batch(pg, device, cornell, small_interior, views, use_graph=1)
    for k in range(4):
        g.render(1, seed=3, max_depth=12)
        # unrelated eager work and synchronisation between replays
        _single(pg, device, cornell, _cam(cornell, 32, 32), 1, 11 + k)
        assert g.film(0)[1] == k + 1
    g_f = [g.film(v)[0] for v in range(len(views))]
    g.close()
    for a, b in zip(e_f, g_f):
        assert np.array_equal(a, b)


def test_chunks_and_autotune_keep_images(pg, device, cornell, small_interior, views):
    one = _batch(pg, device, cornell, small_interior, views, use_graph=1)
    one.render(2, seed=5, max_depth=12)
    f1 = [one.film(v)[0] for v in range(len(views))]
    one.close()
    # at most one view per chunk
    many = _batch(pg, device, cornell, small_interior, views, use_graph=1, queue_paths=1,
                  trace_blocks_per_cu=2, shade_blocks_per_cu=4, stack=64)
    many.render(2, seed=5, max_depth=12)
    assert many.stats()["chunks"] == len(views)
    f2 = [many.film(v)[0] for v in range(len(views))]
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b)
    # persistent lanes with ray claims
    many.configure(use_graph=1, dyn_fetch=1)
    many.clear()
    many.render(2, seed=5, max_depth=12)
    for v, a in enumerate(f1):
        assert np.array_equal(a, many.film(v)[0])
    best, ms = many.autotune(seed=5, max_depth=12)
    assert ms > 0 and best["stack"] >= 16
    many.render(2, seed=5, max_depth=12)
    f3 = [many.film(v)[0] for v in range(len(views))]
    many.close()
    for a, b in zip(f1, f3):
        assert np.array_equal(a, b)


def test_batch_errors(pg, device, cornell):
    b = pg.Batch(device)
    with pytest.raises(pg.PgError):
        b.render(1)  # no views
    b.add_scene(cornell)
    with pytest.raises(pg.PgError):
        b.add_view(3, _cam(cornell, 8, 8))
    b.add_view(0, _cam(cornell, 8, 8))
    with pytest.raises(pg.PgError):
        b.render(1, integrator=2)
    with pytest.raises(pg.PgError):
        b.configure(stack=65)
    b.close()
