"""GPU wavefront renderer vs the CPU oracle (NEE + MIS path tracer, Renderer.h:269-354,
531-618; film Imaging.h:186-272).

* Same random stream on both sides (counter RNG keyed by seed, sample, pixel, dim):
  the films agree pixel by pixel except where a few-ulp difference in a
  transcendental (device libm vs glibc) flips a discrete decision on a path; the
  pixels that differ pass a sign test per primary-hit class (tests/exchangeable.py:
  a bias of either side in one BSDF makes them lean to one sign).
* Statistical bar: the GPU image's relMSE against the FP64 4096-SPP oracle image
  lies within the mean +- 3 sigma of K=8 equal-SPP oracle runs.
* Closed form: white furnace (convex diffuse object under a constant environment).
* Unbiasedness of the area-light NEE + MIS against an FP64 estimator that only samples
  the BSDF (tests/independent.py).
"""
import json

import numpy as np
import pytest

import exchangeable as ex
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _render_gpu(pg, device, data, spp, seed, max_depth=8):
    sc = pg.Scene(device, data)
    rt = pg.RayTracer(sc, seed=seed, max_depth=max_depth)
    rt.clear()
    rt.render(spp)
    film, n = rt.film()
    assert n == spp
    return film, rt


def _pixel_agreement(a, b):
    """Fraction of pixels whose three channels agree to 1e-3 relative."""
    rel = np.abs(a - b) / (np.abs(b) + 1e-3)
    return float(np.mean(np.all(rel < 1e-3, axis=-1)))


def test_cornell_same_stream_matches_oracle(pg, device, cornell):
    from oracle.pg_oracle import OracleScene
    film, _ = _render_gpu(pg, device, cornell, spp=8, seed=5)
    osc = OracleScene(cornell)
    ref, _ = osc.render(8, seed=5, accum_f64=False, env_mode=1)
    agree = _pixel_agreement(film, ref.astype(np.float32))
    assert agree > 0.97, agree
    # and the images are statistically indistinguishable
    assert pg.relmse(film / 8, ref / 8) < 1e-3
    ex.assert_exchangeable(film, ref, ex.primary_classes(pg, pg.Scene(device, cornell), cornell))


def test_interior_same_stream_matches_oracle(pg, device, small_interior):
    """Environment-lit interior: alias-table env sampling, conductors, glass."""
    from oracle.pg_oracle import OracleScene
    film, _ = _render_gpu(pg, device, small_interior, spp=2, seed=3, max_depth=12)
    osc = OracleScene(small_interior)
    ref, _ = osc.render(2, seed=3, max_depth=12, accum_f64=False, env_mode=1)
    agree = _pixel_agreement(film, ref.astype(np.float32))
    assert agree > 0.90, agree
    ex.assert_exchangeable(film, ref, ex.primary_classes(pg, pg.Scene(device, small_interior), small_interior))


@pytest.mark.parametrize("env_visibility", [0, 1])
def test_window_room_same_stream_matches_oracle(pg, device, env_visibility):
    """Closed room lit through one window, both environment shadow-ray targets: the
    reference's scene-bound point (Renderer.h:331-332) and the light direction."""
    from dataclasses import replace
    from oracle.pg_oracle import OracleScene
    base = pg.generate_scene(101, 0)
    data = replace(base, camera=dict(base.camera, width=180, height=320))
    sc = pg.Scene(device, data)
    rt = pg.RayTracer(sc, seed=4, max_depth=12, env_visibility=env_visibility)
    rt.clear()
    rt.render(4)
    film, _ = rt.film()
    ref, _ = OracleScene(data).render(4, seed=4, max_depth=12, accum_f64=False, env_mode=1,
                                      env_visibility=env_visibility)
    agree = _pixel_agreement(film, ref.astype(np.float32))
    assert agree > 0.90, agree
    ex.assert_exchangeable(film, ref, ex.primary_classes(pg, sc, data))


def test_cornell_relmse_within_reference_band(pg, device, cornell):
    band = json.load(open(GOLDEN / "relmse_band_config1.json"))
    ref = np.load(GOLDEN / "oracle_config1_4096spp.npz")["image"]
    spp = band["spp"]
    film, _ = _render_gpu(pg, device, cornell, spp=spp, seed=1000)
    r = pg.relmse(film / spp, ref)
    lo, hi = band["mean"] - 3 * band["std"], band["mean"] + 3 * band["std"]
    assert lo <= r <= hi, (r, lo, hi)


def test_area_light_nee_and_mis_are_unbiased(pg, device):
    """The device's area-light next-event estimation and MIS-weighted emission
    (kernels.hip next_event / k_shade; Renderer.h:269-310, 536-555) against the FP64
    numpy estimator that samples only the BSDF (tests/independent.py), on the closed
    room with two area lights and an occluder at max depth 64: every 8 x 8 pixel block
    sum within 4.5 standard errors, the image total within 1%."""
    import independent as ind

    def render(data, seed):
        sc = pg.Scene(device, data)
        rt = pg.RayTracer(sc, seed=seed, max_depth=64)
        rt.clear()
        rt.render(256)
        film, n = rt.film()
        assert n == 256
        return ind.block_sums(film / 256)

    z, total = ind.against_fixture(pg, render)
    assert np.abs(z).max() < 4.5, np.round(z, 2)
    assert abs(total - 1.0) < 0.01, total


def test_render_is_deterministic(pg, device, cornell):
    a, _ = _render_gpu(pg, device, cornell, spp=2, seed=9)
    b, _ = _render_gpu(pg, device, cornell, spp=2, seed=9)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _icosphere(pg, radius, centre):
    d = pg.generate_scene(100, 0)
    # a frequency-4 icosphere is the last 320 triangles of the test interior
    pos = d.positions[-320:].astype(np.float64)
    c0 = pos.reshape(-1, 3).mean(0)
    r0 = np.linalg.norm(pos.reshape(-1, 3) - c0, axis=1).mean()
    pos = (pos - c0) / r0 * radius + np.asarray(centre)
    nrm = (pos - np.asarray(centre)) / radius
    return pos.astype(np.float32), nrm.astype(np.float32)


def test_white_furnace(pg, device):
    """Convex diffuse sphere (albedo 0.5) under a constant environment Le = 1: the
    surface converges to 0.5 and the background to 1 (per pixel the box-filter
    rounding varies how many samples land, so the check is on the means).

    The reference aims environment shadow rays at sceneCentre + wi * sceneRadius
    (Renderer.h:331), which self-occludes on a scene as compact as one sphere and
    darkens it; two tiny far-away triangles widen the scene bounds so that the
    shadow rays leave along wi and the closed form applies."""
    from pathguidingdissertation_amd.scene import MATERIAL_DTYPE, SceneData
    pos, nrm = _icosphere(pg, 1.0, (0.0, 0.0, 0.0))
    far = np.array([[[1e4, 1e4, 1e4], [1e4 + 1e-2, 1e4, 1e4], [1e4, 1e4 + 1e-2, 1e4]],
                    [[-1e4, -1e4, -1e4], [-1e4 + 1e-2, -1e4, -1e4], [-1e4, -1e4 + 1e-2, -1e4]]], np.float32)
    pos = np.concatenate([pos, far])
    nrm = np.concatenate([nrm, np.tile(np.array([0, 0, 1], np.float32), (2, 3, 1))])
    mats = np.zeros(1, MATERIAL_DTYPE)
    mats["type"] = 0
    mats["albedo"] = 0.5
    env = np.ones((64, 128, 3), np.float32)
    cam = {"from": [0.0, 0.0, 4.0], "to": [0.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0], "fov": 40.0,
           "width": 96, "height": 96}
    data = SceneData(pos, nrm, np.zeros(pos.shape[0], np.uint32), mats, env, cam)
    film, rt = _render_gpu(pg, device, data, spp=64, seed=2)
    img = film / 64
    sc = pg.Scene(device, data)
    rays = sc.generate_rays(np.stack(np.meshgrid(np.arange(96) + 0.5, np.arange(96) + 0.5), -1).reshape(-1, 2))
    ids, *_ = sc.traverse(rays)
    hit = (ids != 0xFFFFFFFF).reshape(96, 96)
    inner = hit & np.roll(hit, 1, 0) & np.roll(hit, -1, 0) & np.roll(hit, 1, 1) & np.roll(hit, -1, 1)
    outer = ~hit & ~np.roll(hit, 1, 0) & ~np.roll(hit, -1, 0) & ~np.roll(hit, 1, 1) & ~np.roll(hit, -1, 1)
    outer[0, :] = False
    outer[:, 0] = False  # row/column 0 receive half the samples (Imaging.h:199-200 rounding)
    assert abs(img[inner].mean() - 0.5) < 0.01, img[inner].mean()
    assert abs(img[outer].mean() - 1.0) < 0.01, img[outer].mean()


@pytest.mark.gpu
def test_ray_sorting_keeps_images(pg, device, small_interior):
    """Tracing secondary and shadow rays in (octant, origin Morton) order, or with other
    launch shapes, changes only the order in which rays are traced: films are identical."""
    sc = pg.Scene(device, small_interior)
    films, counts = [], []
    for kw in (dict(), dict(sort_rays=1), dict(sort_rays=1, trace_blocks_per_cu=4, shade_blocks_per_cu=16),
               dict(dyn_fetch=1), dict(dyn_fetch=1, trace_blocks_per_cu=3)):
        rt = pg.RayTracer(sc, seed=21, max_depth=12, **kw)
        rt.set_counting(True)
        rt.reset_stats()
        rt.clear()
        rt.render(3)
        films.append(rt.film()[0])
        st = rt.stats()
        counts.append(tuple(st[k] for k in ("closest_rays", "shadow_rays", "nodes_visited", "tris_tested",
                                             "shadow_nodes_visited", "shadow_tris_tested")))
        rt.set_counting(False)
    for f, c in zip(films[1:], counts[1:]):
        assert np.array_equal(films[0], f)
        assert c == counts[0]


@pytest.mark.gpu
def test_lds_stack_split_keeps_images(pg, device):
    """Traversal stacks with 16, 32 or 64 entries per lane in LDS and the deeper ones in
    global memory: the ~1.09M-triangle interior's BVH is 26 levels deep, so with 16
    entries in LDS the deep pushes and pops go to global memory. Films and traversal
    counts are identical for every split, with and without persistent lanes; other
    values are refused."""
    import dataclasses
    d = pg.generate_scene(2, 0)
    d = dataclasses.replace(d, camera=dict(d.camera, width=320, height=180))
    sc = pg.Scene(device, d)
    assert sc.bvh_info()["max_depth"] > 16
    films, counts = [], []
    for kw in (dict(stack_lds=64), dict(stack_lds=16), dict(stack_lds=32), dict(stack_lds=16, dyn_fetch=1),
               dict(stack_lds=16, trace_blocks_per_cu=12), dict(stack_lds=0, dyn_fetch=1, trace_blocks_per_cu=3)):
        rt = pg.RayTracer(sc, seed=5, max_depth=12, **kw)
        rt.set_counting(True)
        rt.reset_stats()
        rt.clear()
        rt.render(4)
        films.append(rt.film()[0])
        st = rt.stats()
        counts.append(tuple(st[k] for k in ("closest_rays", "shadow_rays", "nodes_visited", "tris_tested",
                                             "shadow_nodes_visited", "shadow_tris_tested")))
        rt.set_counting(False)
    for f, c in zip(films[1:], counts[1:]):
        assert np.array_equal(films[0].view(np.uint32), f.view(np.uint32))
        assert c == counts[0]
    assert films[0].sum() > 0
    for bad in (8, 80):
        rt = pg.RayTracer(sc, seed=5, max_depth=12, stack_lds=bad)
        with pytest.raises(pg.PgError):
            rt.render(1)


@pytest.mark.gpu
def test_samples_per_pass_keep_images(pg, device, small_interior):
    """Several samples traced together in one wavefront pass (samples_per_pass) change
    only which rays share a launch: films are bit-identical to one sample per pass, also
    when the SPP count is not a multiple of it and across successive render calls."""
    sc = pg.Scene(device, small_interior)
    films = {}
    for spp_pass in (1, 3, 0):
        rt = pg.RayTracer(sc, seed=9, max_depth=12, samples_per_pass=spp_pass)
        rt.clear()
        rt.render(5)
        rt.render(2)
        film, n = rt.film()
        st = rt.stats()
        assert n == 7
        films[spp_pass] = film
    assert np.array_equal(films[1], films[3]) and np.array_equal(films[1], films[0])
    assert films[1].sum() > 0


def test_automatic_samples_per_pass_takes_the_whole_call(pg, device, small_interior):
    """samples_per_pass = 0 traces a call's 70 samples in one pass on a small image (well
    under 2^26 paths); the film equals one sample per pass and 16 per pass bit for bit."""
    sc = pg.Scene(device, small_interior)
    films = []
    for spp_pass in (1, 16, 0):
        rt = pg.RayTracer(sc, seed=13, max_depth=12, samples_per_pass=spp_pass)
        rt.clear()
        rt.render(70)
        film, n = rt.film()
        assert n == 70
        films.append(film)
    assert np.array_equal(films[0], films[1]) and np.array_equal(films[0], films[2])


def test_stale_scene_raises_instead_of_rendering_another(pg, device, cornell):
    """A device context holds one uploaded scene: a Scene (and its RayTracers) created
    before another Scene on the same Device raise instead of silently rendering or
    tracing the other scene; creating it again makes it current."""
    a = pg.Scene(device, cornell)
    rt = pg.RayTracer(a, seed=1)
    rt.clear()
    rt.render(1)
    pg.Scene(device, pg.generate_scene(100, 0))
    with pytest.raises(RuntimeError):
        rt.render(1)
    with pytest.raises(RuntimeError):
        a.traverse(np.zeros((1, 6), np.float32))
    b = pg.Scene(device, cornell)
    rt2 = pg.RayTracer(b, seed=1)
    rt2.clear()
    rt2.render(1)
    assert rt2.getSPP() == 1
