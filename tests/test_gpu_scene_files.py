"""Scenes read from files on the GPU: albedo textures, Oren-Nayar, Plastic and
Layered BSDFs, the device tonemap and the image writers, against the CPU oracle.

* Same random stream on both sides: the films agree pixel by pixel except where a
  few-ulp difference in a transcendental (device libm vs glibc: acosf, cosf, sinf,
  tanf, atan2f, powf) flips a discrete decision on a path.
* The reference's Cornell box as the reference's own loader reads it
  (tests/golden/ref_loader_cornell.npz) renders like the oracle on the same stream.
* Film::tonemap on the device equals the oracle's to one 8-bit level (powf ulps).
"""
import numpy as np
import pytest

import exchangeable as ex
from conftest import GOLDEN, make_scene_dir

pytestmark = pytest.mark.gpu


def _agreement(a, b, tol=1e-3):
    rel = np.abs(a - b) / (np.abs(b) + 1e-3)
    return float(np.mean(np.all(rel < tol, axis=-1)))


def _render(pg, device, data, spp, seed, max_depth=8, **kw):
    sc = pg.Scene(device, data)
    rt = pg.RayTracer(sc, seed=seed, max_depth=max_depth, **kw)
    rt.clear()
    rt.render(spp)
    film, n = rt.film()
    assert n == spp
    return film, rt


def test_textured_scene_same_stream_matches_oracle(pg, device, tmp_path):
    from oracle.pg_oracle import OracleScene
    data = pg.load_scene(make_scene_dir(pg, tmp_path / "scene", width=96, height=64))
    print(f"textured scene: {data.n_tris} tris, {len(data.textures)} textures", flush=True)
    film, _ = _render(pg, device, data, spp=4, seed=3)
    ref, _ = OracleScene(data).render(4, seed=3, max_depth=8, accum_f64=False, env_mode=1)
    agree = _agreement(film, ref.astype(np.float32))
    assert agree > 0.93, agree
    assert pg.relmse(film / 4, ref / 4) < 0.02
    ex.assert_exchangeable(film, ref, ex.primary_classes(pg, pg.Scene(device, data), data))
    # the textures matter: the same scene with constant albedos renders differently
    flat = pg.SceneData(data.positions, data.normals, data.material_index, data.materials, data.env, data.camera)
    film2, _ = _render(pg, device, flat, spp=4, seed=3)
    assert _agreement(film, film2) < 0.9


def test_each_new_bsdf_matches_oracle(pg, device, tmp_path):
    """One material at a time on the floor of the test scene (Oren-Nayar, Plastic
    smooth and rough, Layered over glass), so a mismatch names its BSDF."""
    from dataclasses import replace
    from oracle.pg_oracle import OracleScene
    base = pg.load_scene(make_scene_dir(pg, tmp_path / "scene", width=64, height=48))
    for kind, rough in ((pg.BSDF_ORENNAYAR, 0.7), (pg.BSDF_PLASTIC, 0.05), (pg.BSDF_PLASTIC, 0.8),
                        (pg.BSDF_GLASS + pg.BSDF_LAYERED, 0.0)):
        mats = base.materials.copy()
        mats["type"][0] = kind
        mats["roughness"][0] = rough
        mats["int_ior"][0], mats["ext_ior"][0] = 1.5, 1.0
        data = replace(base, materials=mats)
        print(f"floor bsdf {kind} roughness {rough}", flush=True)
        film, _ = _render(pg, device, data, spp=2, seed=11)
        ref, _ = OracleScene(data).render(2, seed=11, max_depth=8, accum_f64=False, env_mode=1)
        agree = _agreement(film, ref.astype(np.float32))
        assert agree > 0.93, (kind, rough, agree)


def test_reference_cornell_from_loader_vectors(pg, device):
    """The reference's cornell-box scene as its own loader reads it (36 triangles, 1x1
    PNG textures, fov 19.5, 1024x1024 cut to 128x128 here) on the GPU and the oracle."""
    from oracle.pg_oracle import OracleScene
    g = np.load(GOLDEN / "ref_loader_cornell.npz")
    mats = np.zeros(g["materials"].shape[0], pg.scene.MATERIAL_DTYPE)
    mats.view(np.float32).reshape(-1, 16)[:] = g["materials"]
    textures, off = [], 0
    for h, w in g["tex_shape"]:
        textures.append(g["tex_rgb"][off: off + h * w].reshape(h, w, 3))
        off += h * w
    cam = g["camera"]
    camera = {"from": cam[0:3].tolist(), "to": cam[3:6].tolist(), "up": cam[6:9].tolist(), "fov": float(cam[9]),
              "width": 128, "height": 128}
    data = pg.SceneData(g["positions"], g["normals"], g["material_index"], mats, None, camera, uv=g["uv"],
                        material_texture=g["material_texture"], textures=textures)
    film, rt = _render(pg, device, data, spp=8, seed=5)
    ref, _ = OracleScene(data).render(8, seed=5, max_depth=8, accum_f64=False, env_mode=1)
    assert film.mean() > 0
    assert _agreement(film, ref.astype(np.float32)) > 0.97


def test_tonemap_and_image_files(pg, device, tmp_path):
    from oracle import pg_oracle as po
    import ctypes as C
    data = pg.load_scene(make_scene_dir(pg, tmp_path / "scene", width=64, height=48))
    film, rt = _render(pg, device, data, spp=3, seed=2)
    for expo, lmax in ((1.0, 10.0), (3.0, 2.0)):
        got = rt.tonemap(expo, lmax)
        want = np.empty_like(got)
        po.lib().orc_tonemap(po.fp(np.ascontiguousarray(film)), film.shape[0] * film.shape[1], 3, expo, lmax,
                             want.ctypes.data_as(C.POINTER(C.c_uint8)))
        diff = np.abs(got.astype(int) - want.astype(int))
        assert diff.max() <= 1 and np.mean(diff == 0) > 0.999, (expo, lmax, diff.max(), np.mean(diff == 0))
    rt.savePNG(tmp_path / "out.png")
    rt.saveHDR(tmp_path / "out.hdr")
    assert np.array_equal(pg.read_image(tmp_path / "out.png"), rt.tonemap().astype(np.float32))
    img = rt.image()
    hb = pg.read_image(tmp_path / "out.hdr")
    assert np.all(np.abs(hb - img) <= img.max(-1, keepdims=True) / 128 + 1e-30)


def test_texture_entry_errors(pg, device, tmp_path):
    import ctypes as C
    lib, ctx = device.lib, device.ctx
    data = pg.load_scene(make_scene_dir(pg, tmp_path / "scene", width=32, height=32))
    pg.Scene(device, data)
    uv, mt, texels, off, tw, th = data.texture_arrays()
    bad = mt.copy()
    bad[0] = len(data.textures)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    code = lib.pg_scene_textures(ctx, fp(uv), ip(bad), bad.shape[0], fp(texels),
                                 off.ctypes.data_as(C.POINTER(C.c_uint64)), ip(tw), ip(th), len(data.textures))
    assert code == pg._lib.PG_ERR_INVALID
    code = lib.pg_scene_textures(ctx, fp(uv), ip(mt), mt.shape[0] + 1, fp(texels),
                                 off.ctypes.data_as(C.POINTER(C.c_uint64)), ip(tw), ip(th), len(data.textures))
    assert code == pg._lib.PG_ERR_INVALID
