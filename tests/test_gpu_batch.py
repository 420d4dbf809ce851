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


def test_samples_per_pass_keep_batch_views(pg, device, cornell, small_interior, views):
    """Several samples per pixel traced together in a batch pass give the views of one
    sample per pass bit for bit: with graph replay and eager passes, with one or several
    chunks, with a shorter last pass (7 SPP in passes of 3 or 4), across render calls,
    unguided and with a scene's learned guide."""
    scenes = [cornell, small_interior]
    guide = (*_trained_guide(pg, device, small_interior, 12), 12)
    for integ in (0, 3):
        def films(**cfg):
            b = _batch(pg, device, cornell, small_interior, views, **cfg)
            if integ == 3:
                b.set_guide(1, guide[0], guide[1], **dict(GUIDE_CFG, seed=guide[2]))
            b.render(5, seed=9, max_depth=12, integrator=integ)
            b.render(2, seed=9, max_depth=12, integrator=integ)
            out = [b.film(v) for v in range(len(views))]
            b.close()
            return out
        ref = films(use_graph=1)
        for cfg in (dict(use_graph=1, samples_per_pass=3), dict(use_graph=0, samples_per_pass=4),
                    dict(use_graph=1, samples_per_pass=2, queue_paths=1)):
            got = films(**cfg)
            for v in range(len(views)):
                assert got[v][1] == ref[v][1] == 7
                assert np.array_equal(got[v][0], ref[v][0]), (integ, cfg, v)
        assert ref[1][0].sum() > 0
    # the guided interior view is the single-view learned render
    assert np.array_equal(ref[1][0], _single_guided(pg, device, scenes[1], views[1][1], guide, 9, 7))
    b = pg.Batch(device)
    b.add_scene(cornell)
    b.add_view(0, _cam(cornell, 8, 8))
    with pytest.raises(pg.PgError):
        b.configure(samples_per_pass=65)
    b.close()


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
    with pytest.raises(pg.PgError):
        b.render(1, guide_prob=1.5)
    b.close()


GUIDE_CFG = dict(seed=3, log2_table=15, res_min=8.0, res_max=256.0)


def _trained_guide(pg, device, data, seed):
    """Online-trained network of one scene (8 guided record passes, one Adam step each)."""
    sc = pg.Scene(device, data)
    device.mlp_init(**dict(GUIDE_CFG, seed=seed))
    rt = pg.RayTracer(sc, seed=seed, max_depth=12, record_nee=1)
    rt.clear()
    rt.render_online(8, steps=1)
    assert device.mlp_info()["diverged"] == 0
    return device.mlp_params()


def _single_guided(pg, device, data, cam, guide, seed, spp, **params):
    sc = pg.Scene(device, dataclasses.replace(data, camera=cam))
    device.mlp_init(**dict(GUIDE_CFG, seed=guide[2]))
    device.mlp_set_params(guide[0], guide[1])
    rt = pg.RayTracer(sc, seed=seed, max_depth=12, integrator=3, **params)
    rt.clear()
    rt.render(spp)
    return rt.film()[0]


def test_batch_learned_guides_equal_single_renders(pg, device, cornell, small_interior, views):
    """Per-scene learned guides in one batch (both scenes' views share a chunk, so each
    bounce runs one inference launch per scene over the mixed queue): every view equals
    the plain learned-guide pass of that scene with the same network, bit for bit, with
    and without graph replay; a scene without a guide renders unguided."""
    scenes = [cornell, small_interior]
    guides = [(*_trained_guide(pg, device, d, s), s) for d, s in ((cornell, 11), (small_interior, 12))]
    for use_graph in (1, 0):
        b = _batch(pg, device, cornell, small_interior, views, use_graph=use_graph)
        with pytest.raises(pg._lib.PgError):
            b.render(1, seed=7, max_depth=12, integrator=3)  # no guide yet
        for sid in (0, 1):
            b.set_guide(sid, guides[sid][0], guides[sid][1], **dict(GUIDE_CFG, seed=guides[sid][2]))
        b.render(2, seed=7, max_depth=12, integrator=3)
        films = [b.film(v)[0] for v in range(len(views))]
        assert b.stats()["chunks"] == 1
        b.close()
        for v, (sid, cam) in enumerate(views):
            ref = _single_guided(pg, device, scenes[sid], cam, guides[sid], 7, 2)
            assert np.array_equal(films[v], ref), f"view {v} (graph {use_graph}) differs from its single render"
    # only the interior guided: the Cornell views equal unguided single renders
    b = _batch(pg, device, cornell, small_interior, views, use_graph=1)
    b.set_guide(1, guides[1][0], guides[1][1], **dict(GUIDE_CFG, seed=guides[1][2]))
    b.render(2, seed=7, max_depth=12, integrator=3)
    f0 = b.film(0)[0]
    f1 = b.film(1)[0]
    b.close()
    assert np.array_equal(f0, _single(pg, device, cornell, views[0][1], 2, 7))
    assert np.array_equal(f1, _single_guided(pg, device, small_interior, views[1][1], guides[1], 7, 2))


def test_batch_guide_depth_and_cap_follow_single_renders(pg, device, cornell, small_interior, views):
    """`guide_depth` and `guide_prob` reach the batch's guided bounces, and a render with
    other values captures its graph afresh: on one batch with graph replay, renders at
    (depth 0, cap 0.5), (depth 2, cap 0.5) and (depth 2, cap 0.25) each equal the single
    learned-guide renders with the same values, bit for bit."""
    guides = [(*_trained_guide(pg, device, d, s), s) for d, s in ((cornell, 11), (small_interior, 12))]
    scenes = [cornell, small_interior]
    b = _batch(pg, device, cornell, small_interior, views, use_graph=1)
    for sid in (0, 1):
        b.set_guide(sid, guides[sid][0], guides[sid][1], **dict(GUIDE_CFG, seed=guides[sid][2]))
    runs = []
    for depth, cap in ((0, 0.5), (2, 0.5), (2, 0.25)):
        b.clear()
        b.render(2, seed=7, max_depth=12, integrator=3, guide_depth=depth, guide_prob=cap)
        assert b.stats()["graph_nodes"] > 0
        runs.append(((depth, cap), [b.film(v)[0] for v in range(len(views))]))
    b.close()
    for (depth, cap), films in runs:
        for v, (sid, cam) in enumerate(views):
            ref = _single_guided(pg, device, scenes[sid], cam, guides[sid], 7, 2, guide_depth=depth, guide_prob=cap)
            assert np.array_equal(films[v], ref), f"view {v} (guide_depth {depth}, guide_prob {cap}) differs"
    assert not np.array_equal(runs[0][1][1], runs[1][1][1])  # the depth limit changes the interior view


def test_batch_of_more_scenes_than_the_lds_table_groups_every_scene(pg, device, cornell):
    """260 copies of the Cornell box in one batch (one 16x16 view each) with networks on
    scenes 0, 1, 257 and 259 only: the inference grouping counts scenes below 256 per block
    in LDS and the others with global atomics, in the same launches; every guided view
    equals the single learned-guide render and every other view the unguided one."""
    guide = (*_trained_guide(pg, device, cornell, 11), 11)
    cam = _cam(cornell, 16, 16)
    b = pg.Batch(device)
    for i in range(260):
        assert b.add_scene(cornell) == i
        b.add_view(i, cam)
    guided = (0, 1, 257, 259)
    for sid in guided:
        b.set_guide(sid, guide[0], guide[1], **dict(GUIDE_CFG, seed=guide[2]))
    b.clear()
    b.render(2, seed=7, max_depth=12, integrator=3)
    films = [b.film(v)[0] for v in range(260)]
    b.close()
    ref_g = _single_guided(pg, device, cornell, cam, guide, 7, 2)
    ref_u = _single(pg, device, cornell, cam, 2, 7)
    assert not np.array_equal(ref_g, ref_u)
    for v in range(260):
        assert np.array_equal(films[v], ref_g if v in guided else ref_u), f"view {v}"


def test_textured_scene_in_batch_equals_single_render(pg, device, cornell, tmp_path):
    """A scene read from files (albedo textures, Oren-Nayar / Plastic / Layered) in a
    batch beside a procedural one: its views equal its single-view renders."""
    from conftest import make_scene_dir
    tex = pg.load_scene(make_scene_dir(pg, tmp_path / "scene", width=64, height=48))
    b = pg.Batch(device)
    assert b.add_scene(cornell) == 0 and b.add_scene(tex) == 1
    vs = [(1, _cam(tex, 64, 48)), (0, _cam(cornell, 40, 40)), (1, _cam(tex, 50, 30, **{"fov": 70.0}))]
    for sid, cam in vs:
        b.add_view(sid, cam)
    b.configure(use_graph=1)
    b.clear()
    b.render(2, seed=4, max_depth=12)
    films = [b.film(v)[0] for v in range(len(vs))]
    b.close()
    for v, (sid, cam) in enumerate(vs):
        ref = _single(pg, device, [cornell, tex][sid], cam, 2, 4)
        assert np.array_equal(films[v], ref), f"view {v}"
