"""Write the golden fixtures under tests/golden/.

Reference vectors (ref_*.npz) come from the reference's own classes, compiled from
/root/reference by oracle/Makefile into oracle/_ref/libpgref.so (this container only).
Oracle images (oracle_*.npz, *_band.json) come from the CPU oracle in FP64.

    python tests/golden/make_golden.py [--ref] [--images] [--scenes] [--nee] [--independent] [--config2] [--signed-zero]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
OUT = ROOT / "tests" / "golden"

MAT_TYPES = {"diffuse": 0, "mirror": 1, "conductor": 2, "glass": 3, "dielectric": 4, "orennayar": 5, "plastic": 6}
LAYERED = 8
# reference scene directories whose every mesh file is present (no .MISSING_LARGE_BLOBS entry)
COMPLETE_SCENES = ("cornell-box", "veach-mis", "veach-bidir", "MaterialsScene", "staircase2", "living-room",
                   "dining-room")
SCENES = Path("/root/reference/Scenes")


def material(kind, albedo=(0.8, 0.6, 0.4), rough=0.3, eta=(1.0, 0.6, 1.5), k=(0.65, 0.85, 1.0), iors=(1.5, 1.0)):
    m = np.zeros(16, np.float32)
    m[0] = MAT_TYPES[kind]
    m[1:4] = albedo
    m[7:10] = eta
    m[10:13] = k
    m[13], m[14] = iors
    m[15] = rough
    return m


def bsdf_cases(rng, n):
    """Shading frames, wo, random numbers and an evaluation direction per case."""
    def unit(v):
        return v / np.linalg.norm(v, axis=-1, keepdims=True)
    nrm = unit(rng.normal(size=(n, 3)))
    wo = unit(rng.normal(size=(n, 3)))
    flip = np.sum(wo * nrm, 1) < 0
    wo[flip] = -wo[flip]
    # a fifth of the cases look from below (exercises two-sided / transmission paths)
    below = rng.random(n) < 0.2
    wo[below] = -wo[below]
    sd = np.zeros((n, 12), np.float32)
    sd[:, 3:6] = wo
    sd[:, 6:9] = nrm
    sd[:, 9:12] = nrm
    u = rng.random((n, 3)).astype(np.float32)
    we = unit(rng.normal(size=(n, 3))).astype(np.float32)
    return sd, u, we


def make_ref():
    from oracle import pg_oracle as po
    import pathguidingdissertation_amd as pg
    assert po.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    R = po.ref()
    import ctypes as C

    # --- BVH topology + hits (Geometry.h:101-130, 175-192, 293-567)
    for cfg, n_rays in ((1, 8192), (100, 4096)):
        data = pg.generate_scene(cfg, 0)
        osc = po.OracleScene(data)
        from conftest import seeded_rays
        rays = seeded_rays(osc, data, n_rays, seed=100 + cfg)
        pos = np.ascontiguousarray(data.positions)
        h = R.ref_scene_create(po.fp(pos), po.fp(np.ascontiguousarray(data.normals)), data.n_tris)
        nn = R.ref_bvh_node_count(h)
        nodes = np.empty((nn, 8), np.float32)
        idx = np.empty(data.n_tris, np.int32)
        R.ref_bvh_export(h, po.fp(nodes), idx.ctypes.data_as(C.POINTER(C.c_int32)))
        ids = np.empty(n_rays, np.uint32)
        t, a, b, g = (np.empty(n_rays, np.float32) for _ in range(4))
        R.ref_trace_closest(h, po.fp(rays), n_rays, po.up(ids), po.fp(t), po.fp(a), po.fp(b), po.fp(g))
        max_t = np.random.default_rng(cfg).uniform(0.01, 20.0, n_rays).astype(np.float32)
        vis = np.empty(n_rays, np.uint8)
        R.ref_trace_shadow(h, po.fp(rays), po.fp(max_t), n_rays, vis.ctypes.data_as(C.POINTER(C.c_uint8)))
        R.ref_scene_destroy(h)
        extra = {}
        if cfg == 1:
            extra = {"nodes": nodes, "tri_index": idx}
        np.savez_compressed(OUT / f"ref_trace_config{cfg}.npz", rays=rays, id=ids, t=t, alpha=a, beta=b, gamma=g,
                            max_t=max_t, visible=vis, n_nodes=np.int64(nn),
                            topology_sha256=np.bytes_(hashlib.sha256(nodes.tobytes() + idx.tobytes()).hexdigest()),
                            **extra)
        print("trace config", cfg, "nodes", nn)

    # full-size topology hash for the ~1.09M-triangle interior
    data = pg.generate_scene(2, 0)
    pos = np.ascontiguousarray(data.positions)
    h = R.ref_scene_create(po.fp(pos), po.fp(np.ascontiguousarray(data.normals)), data.n_tris)
    nn = R.ref_bvh_node_count(h)
    nodes = np.empty((nn, 8), np.float32)
    idx = np.empty(data.n_tris, np.int32)
    R.ref_bvh_export(h, po.fp(nodes), idx.ctypes.data_as(C.POINTER(C.c_int32)))
    R.ref_scene_destroy(h)
    scene_sha = hashlib.sha256(pos.tobytes() + np.ascontiguousarray(data.normals).tobytes()).hexdigest()
    json.dump({"n_tris": int(data.n_tris), "n_nodes": int(nn), "scene_sha256": scene_sha,
               "topology_sha256": hashlib.sha256(nodes.tobytes() + idx.tobytes()).hexdigest()},
              open(OUT / "ref_topology_config2.json", "w"), indent=1)
    print("config 2 nodes", nn)

    # --- camera (Scene.h:25-57, MyMath.h:286-364)
    cams = []
    for cfg in (1, 100):
        data = pg.generate_scene(cfg, 0)
        rng = np.random.default_rng(cfg)
        pxy = np.stack([rng.uniform(0, data.width, 2048), rng.uniform(0, data.height, 2048)], 1).astype(np.float32)
        ip, cm, rays = np.empty(16, np.float32), np.empty(16, np.float32), np.empty((2048, 6), np.float32)
        R.ref_camera(po.fp(data.camera_array()), po.fp(pxy), 2048, po.fp(ip), po.fp(cm), po.fp(rays))
        cams.append((data.camera_array(), pxy, ip, cm, rays))
    np.savez_compressed(OUT / "ref_camera.npz", **{f"{k}{i}": v for i, c in enumerate(cams)
                                                  for k, v in zip(("cam", "pxy", "inv_proj", "camera", "rays"), c)})

    # --- BSDFs (Materials.h:36-639) with injected random numbers (TestSampler pattern)
    rng = np.random.default_rng(42)
    mats, sds, us, wes, outs = [], [], [], [], []
    for kind, rough in (("diffuse", 0.0), ("mirror", 0.0), ("conductor", 0.3), ("conductor", 0.0),
                        ("conductor", 0.8), ("glass", 0.0), ("dielectric", 0.2), ("dielectric", 0.0)):
        m = material(kind, rough=rough)
        sd, u, we = bsdf_cases(rng, 512)
        for i in range(sd.shape[0]):
            out = np.empty(14, np.float32)
            R.ref_bsdf(po.fp(m), po.fp(sd[i]), po.fp(u[i]), po.fp(we[i]), po.fp(out))
            mats.append(m); sds.append(sd[i]); us.append(u[i]); wes.append(we[i]); outs.append(out)
    np.savez_compressed(OUT / "ref_bsdf.npz", mat=np.array(mats), sd=np.array(sds), u=np.array(us),
                        wi_eval=np.array(wes), out=np.array(outs))
    # Fresnel / GGX helpers
    c = np.linspace(-1.0, 1.0, 257).astype(np.float32)
    fd = np.array([R.ref_fresnel_dielectric(float(x), 1.5, 1.0) for x in c], np.float32)
    eta, k = np.array([0.2, 0.9, 1.1], np.float32), np.array([3.9, 2.4, 2.1], np.float32)
    fc = np.empty((c.size, 3), np.float32)
    for i, x in enumerate(c):
        R.ref_fresnel_conductor(float(x), po.fp(eta), po.fp(k), po.fp(fc[i]))
    hs = rng.normal(size=(256, 3)).astype(np.float32)
    hs /= np.linalg.norm(hs, axis=1, keepdims=True)
    alphas = rng.uniform(0.05, 1.5, 256).astype(np.float32)
    dg = np.array([R.ref_ggx_d(po.fp(hs[i]), float(alphas[i])) for i in range(256)], np.float32)
    wi2 = rng.normal(size=(256, 3)).astype(np.float32)
    wi2 /= np.linalg.norm(wi2, axis=1, keepdims=True)
    gg = np.array([R.ref_ggx_g(po.fp(wi2[i]), po.fp(hs[i]), float(alphas[i])) for i in range(256)], np.float32)
    np.savez_compressed(OUT / "ref_fresnel_ggx.npz", cos=c, fresnel_dielectric=fd, eta=eta, k=k,
                        fresnel_conductor=fc, h=hs, alpha=alphas, ggx_d=dg, wi=wi2, ggx_g=gg)

    # --- environment map (Lights.h:128-228, Sampling.h:105-209) on a small synthetic map
    env = synthetic_env(64, 32)
    he = R.ref_env_create(po.fp(env), 64, 32)
    u2 = rng.random((4096, 2)).astype(np.float32)
    samp = np.empty((4096, 7), np.float32)
    for i in range(4096):
        R.ref_env_sample(he, float(u2[i, 0]), float(u2[i, 1]), po.fp(samp[i]))
    dirs = rng.normal(size=(4096, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ev = np.empty((4096, 4), np.float32)
    for i in range(4096):
        R.ref_env_eval(he, po.fp(dirs[i]), po.fp(ev[i]))
    power = R.ref_env_power(he)
    R.ref_env_destroy(he)
    np.savez_compressed(OUT / "ref_env.npz", env=env, u=u2, sample=samp, dirs=dirs, eval=ev, power=np.float32(power))
    # --- Metrics.h:8-69 on odd-sized images with a wide dynamic range
    w, h = 37, 23
    truth = (rng.random((h, w, 3)) ** 4 * 50).astype(np.float32)
    truth[0, 0] = [5e4, 0.0, 1e-6]
    test = (truth + rng.normal(0, 0.3, truth.shape) * (truth + 0.05)).astype(np.float32)
    test[5, 7] = [1e3, 0.0, -2.0]
    out = np.empty(3, np.float32)
    R.ref_metrics(po.fp(test.reshape(-1)), po.fp(truth.reshape(-1)), test.size, truth.size, w, h, po.fp(out))
    bad = np.empty(3, np.float32)
    R.ref_metrics(po.fp(test.reshape(-1)), po.fp(truth.reshape(-1)), test.size, truth.size - 3, w, h, po.fp(bad))
    np.savez_compressed(OUT / "ref_metrics.npz", test=test, truth=truth, out=out, mismatch=bad)
    print("reference vectors written")


def make_scenes():
    """Vectors for the scene-file path: the Oren-Nayar / Plastic / Layered BSDFs,
    Texture::sample, Film::tonemap, and the reference's loading of its complete
    scene directories (array hashes; the scene files themselves are not copied)."""
    from oracle import pg_oracle as po
    import ctypes as C
    assert po.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    R = po.ref()
    rng = np.random.default_rng(4242)
    # --- BSDFs (Materials.h:641-897) with injected random numbers
    cases = [("orennayar", dict(rough=0.0)), ("orennayar", dict(rough=0.35)), ("orennayar", dict(rough=1.0)),
             ("plastic", dict(rough=0.05, iors=(1.5, 1.0))), ("plastic", dict(rough=0.4, iors=(1.33, 1.0))),
             ("plastic", dict(rough=1.0, iors=(1.6, 1.1))), ("plastic", dict(rough=1e-7, iors=(1.5, 1.0)))]
    layered = [("conductor", dict(rough=0.3)), ("glass", dict(rough=0.0)), ("diffuse", dict(rough=0.0)),
               ("plastic", dict(rough=0.2, iors=(1.5, 1.0)))]
    mats, sds, us, wes, outs = [], [], [], [], []
    for kind, kw in cases + [(k, dict(kw, layered=True)) for k, kw in layered]:
        lay = kw.pop("layered", False)
        m = material(kind, **kw)
        if lay:
            m[0] += LAYERED
        sd, u, we = bsdf_cases(rng, 384)
        for i in range(sd.shape[0]):
            out = np.empty(15, np.float32)
            R.ref_bsdf(po.fp(m), po.fp(sd[i]), po.fp(u[i]), po.fp(we[i]), po.fp(out))
            mats.append(m); sds.append(sd[i]); us.append(u[i]); wes.append(we[i]); outs.append(out)
    np.savez_compressed(OUT / "ref_bsdf_ext.npz", mat=np.array(mats), sd=np.array(sds), u=np.array(us),
                        wi_eval=np.array(wes), out=np.array(outs))
    # --- Texture::sample (Imaging.h:82-108): odd sizes, uv beyond [0,1] and negative
    tex = {}
    for i, (w, h) in enumerate(((1, 1), (5, 3), (16, 9))):
        t = rng.random((h, w, 3)).astype(np.float32)
        tuv = np.concatenate([rng.uniform(-2.5, 3.5, (500, 2)), rng.random((500, 2)),
                              [[0, 0], [1, 1], [0.999999, 0.5], [-0.0, 1.0], [2.0, -3.0]]]).astype(np.float32)
        out = np.empty((tuv.shape[0], 3), np.float32)
        R.ref_texture_sample(po.fp(t), w, h, po.fp(tuv), tuv.shape[0], po.fp(out))
        tex[f"tex{i}"], tex[f"uv{i}"], tex[f"out{i}"] = t, tuv, out
    np.savez_compressed(OUT / "ref_texture.npz", **tex)
    # --- Film::tonemap (Imaging.h:220-244) on accumulated sums
    n = 4096
    rgb = (rng.random((n, 3)) ** 3 * 40.0).astype(np.float32)
    rgb[:16] = 0.0
    rgb[16:32] *= 1e-4
    outs = {}
    for spp, expo, lmax in ((1, 1.0, 10.0), (7, 1.0, 10.0), (64, 2.5, 4.0)):
        o = np.empty((n, 3), np.uint8)
        R.ref_tonemap(po.fp(rgb), n, spp, expo, lmax, o.ctypes.data_as(C.POINTER(C.c_uint8)))
        outs[f"spp{spp}"] = o
    np.savez_compressed(OUT / "ref_tonemap.npz", rgb=rgb, params=np.array([[1, 1.0, 10.0], [7, 1.0, 10.0],
                                                                           [64, 2.5, 4.0]], np.float32), **outs)
    # --- loadScene (SceneLoader.h:104-269) of the complete scene directories
    summary = {}
    jpeg_samples = {}
    for name in COMPLETE_SCENES:
        d = SCENES / name
        h = R.ref_load_scene(str(d).encode())
        nm, ntx = C.c_int(), C.c_int()
        nt = R.ref_loaded_sizes(h, C.byref(nm), C.byref(ntx))
        pos = np.empty((nt, 3, 3), np.float32); nrm = np.empty((nt, 3, 3), np.float32)
        uv = np.empty((nt, 3, 2), np.float32); mi = np.empty(nt, np.uint32)
        mm = np.empty((nm.value, 16), np.float32); mt = np.empty(nm.value, np.int32); cam = np.empty(12, np.float32)
        R.ref_loaded_copy(h, po.fp(pos), po.fp(nrm), po.fp(uv), po.up(mi), po.fp(mm),
                          mt.ctypes.data_as(C.POINTER(C.c_int32)), po.fp(cam))
        names = [R.ref_loaded_texture_name(h, i).decode() for i in range(ntx.value)]
        R.ref_loaded_free(h)
        texs = {}
        for tn in names:
            w, hh = C.c_int(), C.c_int()
            R.ref_texture_load(tn.encode(), C.byref(w), C.byref(hh), None)
            if w.value <= 0 or hh.value <= 0:
                continue
            buf = np.empty((hh.value, w.value, 3), np.float32)
            R.ref_texture_load(tn.encode(), C.byref(w), C.byref(hh), po.fp(buf))
            if tn.lower().endswith((".png", ".hdr")):
                texs[Path(tn).name] = {"shape": [hh.value, w.value],
                                       "sha256": hashlib.sha256(buf.tobytes()).hexdigest()}
            elif tn.lower().endswith((".jpg", ".jpeg")):
                # JPEG decoders may differ by a few levels (the IDCT and colour conversion
                # are not bit-specified): 512 sampled pixels of the reference's decode
                srng = np.random.default_rng(len(jpeg_samples) + 1)
                ys, xs = srng.integers(0, hh.value, 512), srng.integers(0, w.value, 512)
                key = f"{name}/{Path(tn).name}"
                jpeg_samples[key + ":yx"] = np.stack([ys, xs], 1).astype(np.int32)
                jpeg_samples[key + ":rgb8"] = np.rint(buf[ys, xs] * 255).astype(np.uint8)
                texs[Path(tn).name] = {"shape": [hh.value, w.value], "jpeg_samples": key}

        def sha(a):
            return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
        summary[name] = {"n_tris": int(nt), "n_materials": int(nm.value), "camera": cam.tolist(),
                         "positions": sha(pos), "normals": sha(nrm), "uv": sha(uv), "material_index": sha(mi),
                         "materials": sha(mm), "material_texture": mt.tolist(),
                         "textures": [Path(t).name for t in names], "decoded": texs}
        if name == "cornell-box":
            tex_rgb, tex_shape = [], []
            for tn in names:
                w, hh = C.c_int(), C.c_int()
                R.ref_texture_load(tn.encode(), C.byref(w), C.byref(hh), None)
                buf = np.empty((hh.value, w.value, 3), np.float32)
                R.ref_texture_load(tn.encode(), C.byref(w), C.byref(hh), po.fp(buf))
                tex_rgb.append(buf.reshape(-1, 3)); tex_shape.append((hh.value, w.value))
            np.savez_compressed(OUT / "ref_loader_cornell.npz", positions=pos, normals=nrm, uv=uv,
                                material_index=mi, materials=mm, material_texture=mt, camera=cam,
                                tex_rgb=np.concatenate(tex_rgb), tex_shape=np.array(tex_shape, np.int32))
        print("scene", name, nt, "tris")
    json.dump(summary, open(OUT / "ref_loader.json", "w"), indent=1)
    np.savez_compressed(OUT / "ref_jpeg_samples.npz", **jpeg_samples)
    print("scene vectors written")


def make_nee():
    """Vectors for the next-event-estimation constituents over the reference's own Scene:
    the light list (Scene.h:93-100, 133-136), AreaLight::sample / evaluate / PDF / normal
    with Triangle::sample (Lights.h:28-73, Geometry.h:139-150), Scene::sampleLight
    (Scene.h:108-119), Scene::calculateShadingData (Scene.h:152-176), Scene::visible
    (Scene.h:139-146) and Film::splat with the box filter (Imaging.h:147-151, 186-217)."""
    from oracle import pg_oracle as po
    import pathguidingdissertation_amd as pg
    from conftest import seeded_rays
    assert po.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    rng = np.random.default_rng(2468)
    out = {}
    scenes = {"cornell": (pg.generate_scene(1, 0), None), "cornell_env": (pg.generate_scene(1, 0), synthetic_env(64, 32)),
              "interior": (pg.generate_scene(100, 0), "own")}
    for key, (data, env) in scenes.items():
        if isinstance(env, str):
            env = data.env
        rs = po.RefFullScene(data, env)
        kinds, tris = rs.lights()
        out[f"{key}:light_kind"], out[f"{key}:light_tri"] = kinds, tris
        if env is not None and key != "interior":
            out[f"{key}:env"] = env
        area = np.nonzero(kinds == 0)[0]
        if area.size:
            u2 = rng.random((256, 2)).astype(np.float32)
            u2[:4] = [[0, 0], [1, 1], [0, 1], [1, 0]]
            wi = rng.normal(size=(256, 3)).astype(np.float32)
            wi /= np.linalg.norm(wi, axis=1, keepdims=True)
            al = np.stack([rs.area_light(int(l), u2[i], wi[i]) for l in area for i in range(256)])
            out[f"{key}:area_u2"], out[f"{key}:area_wi"], out[f"{key}:area_out"] = u2, wi, al
        u = np.concatenate([rng.random(4093), [0.0, 0.99999994, 1.0]]).astype(np.float32)
        idx, pmf = rs.sample_light(u)
        out[f"{key}:sl_u"], out[f"{key}:sl_idx"], out[f"{key}:sl_pmf"] = u, idx, pmf
        osc = po.OracleScene(data)
        rays = seeded_rays(osc, data, 2048, seed=77)
        out[f"{key}:rays"], out[f"{key}:shading"] = rays, rs.shading_data(rays)
        sd = out[f"{key}:shading"]
        hit = sd[:, 14] < np.finfo(np.float32).max
        lo, hi = data.positions.reshape(-1, 3).min(0), data.positions.reshape(-1, 3).max(0)
        p1 = np.concatenate([sd[hit, 0:3][:768], rng.uniform(lo, hi, (768, 3))]).astype(np.float32)
        p2 = np.concatenate([rng.uniform(lo, hi, (p1.shape[0] - 512, 3)), sd[hit, 0:3][-512:]]).astype(np.float32)
        out[f"{key}:vis_p1"], out[f"{key}:vis_p2"], out[f"{key}:visible"] = p1, p2, rs.visible(p1, p2)
        del rs
    w, h = 7, 5
    xy = np.concatenate([rng.uniform(-1.5, [w + 1.5, h + 1.5], (3000, 2)),
                         np.array([[0.5, 0.5], [1.5, 2.5], [-0.5, 0.0], [6.5, 4.5], [7.0, 5.0], [0.0, 0.0],
                                   [3.4999998, 2.5000002]])]).astype(np.float32)
    L = (rng.random((xy.shape[0], 3)) * 10).astype(np.float32)
    out["splat:wh"], out["splat:xy"], out["splat:L"] = np.array([w, h], np.int32), xy, L
    out["splat:film"] = po.ref_film_splat(w, h, xy, L)
    np.savez_compressed(OUT / "ref_nee.npz", **out)
    print("next-event vectors written")


def make_independent(spp=4096, seed=2024):
    """Block sums of the area-light room (tests/independent.py) from the FP64 numpy
    BSDF-sampling estimator, which shares no code with any next-event estimation, with
    their standard errors: the reference point of the NEE + MIS unbiasedness tests."""
    from oracle import pg_oracle as po
    import pathguidingdissertation_amd as pg
    import independent as ind
    data = ind.area_light_box(pg)
    osc = po.OracleScene(data)
    mean, se = ind.bsdf_only_estimate(data, osc.generate_rays, spp, seed)
    np.savez_compressed(OUT / "independent_area_box.npz", block_sum=mean, stderr=se, spp=spp, seed=seed,
                        block=ind.BLOCK, width=data.width, height=data.height)
    print("independent estimate: relative standard error max", float((se / mean).max()))


CONFIG2_CROPS = ((560, 420, 64, 64), (200, 150, 64, 64), (880, 330, 64, 64))


def make_config2(spp_ref=4096, k=8, spp=64):
    """Three 64 x 64 crops of the ~1.09M-triangle interior at 1280x720, max depth 12
    (objects, a sunlit wall, the far corner): the FP64 oracle at 4096 SPP and the
    relMSE of k equal-SPP oracle runs against it (the band a renderer's crops must
    fall in). Only the crops are traced; each pixel keeps its own random stream."""
    from oracle import pg_oracle as po
    import pathguidingdissertation_amd as pg
    data = pg.generate_scene(2, 0)
    osc = po.OracleScene(data)

    def crops(film):
        return np.stack([film[y:y + h, x:x + w] for x, y, w, h in CONFIG2_CROPS])

    ref = None
    for c in CONFIG2_CROPS:
        film, _ = osc.render(spp_ref, seed=9002, max_depth=12, accum_f64=True, crop=c)
        ref = film if ref is None else ref + film
    ref_c = crops(ref / spp_ref)
    inner = np.s_[:, 1:-1, 1:-1]  # the box filter's border pixels also receive outside samples
    vals, trimmed = [], []
    for s in range(1, k + 1):
        img = None
        for c in CONFIG2_CROPS:
            film, _ = osc.render(spp, seed=s, max_depth=12, accum_f64=True, crop=c)
            img = film if img is None else img + film
        img_c = crops(img / spp)
        e = ((img_c[inner] - ref_c[inner]) ** 2 / (ref_c[inner] ** 2 + 1e-2)).mean(-1).reshape(-1)
        vals.append(float(e.mean()))
        trimmed.append(float(np.sort(e)[: e.size - e.size // 1000].mean()))  # without the worst 0.1%
    np.savez_compressed(OUT / "oracle_config2_crops.npz", image=ref_c.astype(np.float32),
                        crops=np.array(CONFIG2_CROPS, np.int32), spp=spp_ref, seed=9002,
                        band_spp=spp, band_relmse=np.array(vals), band_trimmed=np.array(trimmed))
    print("config 2 crops: band", np.mean(vals), np.std(vals, ddof=1), vals, "trimmed", trimmed)


CONFIG3_CROPS = ((300, 560, 64, 64), (160, 400, 64, 64), (480, 880, 64, 64))


def make_config3(spp_ref=4096, k=8, spp=64):
    """Three 64 x 64 crops of the window-lit room (config 3, 720 x 1280, max depth 12,
    environment shadow rays along the light direction, as the guided comparisons run):
    the sun patch below the window, the indirectly lit wall above it and a dark corner
    lit only through several bounces. The FP64 oracle at 4096 SPP and the relMSE of k
    equal-SPP oracle runs against it, as make_config2."""
    from oracle import pg_oracle as po
    import pathguidingdissertation_amd as pg
    data = pg.generate_scene(3, 0)
    osc = po.OracleScene(data)

    def crops(film):
        return np.stack([film[y:y + h, x:x + w] for x, y, w, h in CONFIG3_CROPS])

    ref = None
    for c in CONFIG3_CROPS:
        film, _ = osc.render(spp_ref, seed=9003, max_depth=12, accum_f64=True, env_visibility=1, crop=c)
        ref = film if ref is None else ref + film
    ref_c = crops(ref / spp_ref)
    inner = np.s_[:, 1:-1, 1:-1]
    vals, trimmed = [], []
    for s in range(1, k + 1):
        img = None
        for c in CONFIG3_CROPS:
            film, _ = osc.render(spp, seed=s, max_depth=12, accum_f64=True, env_visibility=1, crop=c)
            img = film if img is None else img + film
        img_c = crops(img / spp)
        e = ((img_c[inner] - ref_c[inner]) ** 2 / (ref_c[inner] ** 2 + 1e-2)).mean(-1).reshape(-1)
        vals.append(float(e.mean()))
        trimmed.append(float(np.sort(e)[: e.size - e.size // 1000].mean()))  # without the worst 0.1%
    np.savez_compressed(OUT / "oracle_config3_crops.npz", image=ref_c.astype(np.float32),
                        crops=np.array(CONFIG3_CROPS, np.int32), spp=spp_ref, seed=9003,
                        band_spp=spp, band_relmse=np.array(vals), band_trimmed=np.array(trimmed))
    print("config 3 crops: band", np.mean(vals), np.std(vals, ddof=1), vals, "trimmed", trimmed)


def signed_zero_positions(n=3000, seed=31):
    """Triangles whose coordinates are often exactly -0.0 or +0.0 (both signs), so node
    bounds meet ties between -0 and +0 (AABB::extend keeps the first one)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 3, 3))
    pos[:, 1:] = pos[:, :1] + rng.uniform(-0.2, 0.2, (n, 2, 3))
    z = rng.random((n, 3, 3)) < 0.3
    pos[z] = np.where(rng.random(z.sum()) < 0.5, -0.0, 0.0)
    half = rng.random(n) < 0.5  # half of the triangles lie on the x >= 0 side only
    pos[half, :, 0] = np.abs(pos[half, :, 0])
    pos[half, :, 0] = np.where(pos[half, :, 0] == 0, np.where(rng.random((half.sum(), 3)) < 0.5, -0.0, 0.0),
                               pos[half, :, 0])
    return np.ascontiguousarray(pos.astype(np.float32))


def make_signed_zero():
    """The reference's BVH (Geometry.h:274-446) over the signed-zero triangles."""
    from oracle import pg_oracle as po
    import ctypes as C
    R = po.ref()
    pos = signed_zero_positions()
    n = pos.shape[0]
    nrm = np.zeros_like(pos)
    nrm[..., 2] = 1.0
    h = R.ref_scene_create(po.fp(pos), po.fp(nrm), n)
    nn = R.ref_bvh_node_count(h)
    nodes = np.empty((nn, 8), np.float32)
    idx = np.empty(n, np.int32)
    R.ref_bvh_export(h, po.fp(nodes), idx.ctypes.data_as(C.POINTER(C.c_int32)))
    R.ref_scene_destroy(h)
    neg0 = int(np.sum(np.signbit(nodes[:, [0, 1, 2, 4, 5, 6]]) & (nodes[:, [0, 1, 2, 4, 5, 6]] == 0)))
    np.savez_compressed(OUT / "ref_bvh_signed_zero.npz", positions=pos, nodes=nodes, tri_index=idx)
    print("signed-zero BVH:", nn, "nodes,", neg0, "bounds that are -0")


def synthetic_env(w, h):
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    th, ph = np.meshgrid(np.pi * v, 2 * np.pi * u, indexing="ij")
    d = np.stack([np.cos(ph) * np.sin(th), np.cos(th), np.sin(ph) * np.sin(th)], -1)
    sky = np.where(d[..., 1:2] > 0, np.array([0.3, 0.5, 1.0]) * (0.2 + d[..., 1:2]), np.array([0.1, 0.08, 0.05]))
    sun = np.array([0.6, 0.7, 0.3])
    sun /= np.linalg.norm(sun)
    lobe = 200.0 * np.exp(-(1 - np.sum(d * sun, -1)) / 0.01)[..., None]
    return np.ascontiguousarray((sky + lobe).astype(np.float32))


def make_images(spp_ref=4096, k=8, spp=16):
    """FP64 oracle image of the Cornell box and the relMSE band of equal-SPP CPU runs."""
    from oracle import pg_oracle as po
    import pathguidingdissertation_amd as pg
    data = pg.generate_scene(1, 0)
    osc = po.OracleScene(data)
    film, st = osc.render(spp_ref, seed=9001, accum_f64=True)
    ref = film / spp_ref
    np.savez_compressed(OUT / "oracle_config1_4096spp.npz", image=ref.astype(np.float32), spp=spp_ref, seed=9001)
    vals = []
    for s in range(1, k + 1):
        f, _ = osc.render(spp, seed=s, accum_f64=False)
        img = (f / spp).astype(np.float64)
        vals.append(float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2))))
    band = {"config": 1, "spp": spp, "k": k, "relmse": vals, "mean": float(np.mean(vals)),
            "std": float(np.std(vals, ddof=1)), "reference_spp": spp_ref, "reference_seed": 9001}
    json.dump(band, open(OUT / "relmse_band_config1.json", "w"), indent=1)
    print("band", band)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", action="store_true")
    ap.add_argument("--images", action="store_true")
    ap.add_argument("--scenes", action="store_true")
    ap.add_argument("--nee", action="store_true")
    ap.add_argument("--independent", action="store_true")
    ap.add_argument("--config2", action="store_true")
    ap.add_argument("--config3", action="store_true")
    ap.add_argument("--signed-zero", action="store_true")
    a = ap.parse_args()
    if a.ref:
        make_ref()
    if a.images:
        make_images()
    if a.scenes:
        make_scenes()
    if a.nee:
        make_nee()
    if a.independent:
        make_independent()
    if a.config2:
        make_config2()
    if a.config3:
        make_config3()
    if a.signed_zero:
        make_signed_zero()
