"""Scene files and the BSDFs they use, on the CPU.

* The oracle's Oren-Nayar, Plastic and Layered BSDFs and Texture::sample against
  vectors written by the reference's own classes (tests/golden/ref_bsdf_ext.npz,
  ref_texture.npz), and Film::tonemap (ref_tonemap.npz).
* The library's loader (pg_scene_load) against the reference's loading of its
  complete scene directories (tests/golden/ref_loader.json: array hashes written by
  GEMScene / GEMModelLoader / Matrix / Triangle / Texture of the reference); this
  part needs the reference's scene files and is skipped where they are absent.
* A scene directory written by the library's own writers (.gem, PNG, HDR) loads
  back to the geometry, materials and textures it was written with.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, make_scene_dir

SCENES = Path("/root/reference/Scenes")


@pytest.fixture(scope="module")
def po():
    from oracle import pg_oracle
    return pg_oracle


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_new_bsdfs_match_reference(po):
    """Sample / evaluate / pdf / invert / flags of Oren-Nayar (sigma 0, 0.35, 1),
    Plastic (four roughness / IOR pairs, one at the exponent clamp) and LayeredBSDF
    over conductor, glass, diffuse and plastic, with injected random numbers. acosf /
    cosf / sinf / tanf / atan2f / powf come from this platform's libm, so results are
    compared to a relative 2e-5 (Oren-Nayar and Phong lobes chain several of them)."""
    g = np.load(GOLDEN / "ref_bsdf_ext.npz")
    L = po.lib()
    import ctypes as C
    n = g["mat"].shape[0]
    for i in range(n):
        m, sd, u, we, ref = (np.ascontiguousarray(g[k][i], np.float32) for k in ("mat", "sd", "u", "wi_eval", "out"))
        wi, f, pdf = np.empty(3, np.float32), np.empty(3, np.float32), np.empty(1, np.float32)
        L.orc_bsdf_sample(po.fp(m), po.fp(sd), po.fp(u), po.fp(wi), po.fp(f), po.fp(pdf))
        fe, pe = np.empty(3, np.float32), np.empty(1, np.float32)
        L.orc_bsdf_eval(po.fp(m), po.fp(sd), po.fp(we), po.fp(fe), po.fp(pe))
        got = np.concatenate([wi, f, pdf, fe, pe])
        exp = ref[:11]
        err = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-6)
        ok = (err <= 2e-5) | (np.abs(got - exp) <= 1e-6) | (np.isnan(got) & np.isnan(exp))
        assert ok.all(), (i, m[0], m[15], got, exp)
        spec, two = C.c_int32(), C.c_int32()
        L.orc_bsdf_flags(po.fp(m), C.byref(spec), C.byref(two))
        assert spec.value == int(ref[13]) and two.value == int(ref[14]), (i, m[0], spec.value, two.value, ref[13:])
        if int(m[0]) % 8 == 5:  # OrenNayarBSDF::invert is the cosine warp's inverse
            uv = np.empty(2, np.float32)
            L.orc_bsdf_invert(po.fp(m), po.fp(sd), po.fp(wi), po.fp(uv))
            assert np.allclose(uv, ref[11:13], rtol=2e-5, atol=2e-6)


def test_new_bsdfs_sample_consistent_with_eval(po):
    """Closed-form checks independent of the vectors: Oren-Nayar's cosine pdf
    integrates to 1 over the hemisphere and sigma 0 is Lambert (A = 1, B = 0); for
    Plastic, the hemisphere integral of the pdf (estimated by uniform directions)
    equals the fraction of its samples that land above the surface (the Phong lobe
    reflected about the half vector can leave the hemisphere, where the pdf is 0)."""
    rng = np.random.default_rng(3)
    L = po.lib()
    sd = np.array([0, 0, 0, 0.3, 0.2, 0.93, 0, 0, 1, 0, 0, 1], np.float32)
    sd[3:6] /= np.linalg.norm(sd[3:6])
    n = 30000
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w[:, 2] = np.abs(w[:, 2])
    w = w.astype(np.float32)
    u = rng.random((n, 3)).astype(np.float32)
    for kind, rough in ((5, 0.0), (5, 0.8), (6, 0.2), (6, 0.9)):
        m = np.zeros(16, np.float32)
        m[0], m[1:4], m[13], m[14], m[15] = kind, (0.8, 0.6, 0.4), 1.5, 1.0, rough
        pdfs = np.empty(n, np.float32)
        fs = np.empty((n, 3), np.float32)
        up = 0
        f, p, wi = np.empty(3, np.float32), np.empty(1, np.float32), np.empty(3, np.float32)
        for i in range(n):
            L.orc_bsdf_eval(po.fp(m), po.fp(sd), po.fp(w[i]), po.fp(f), po.fp(p))
            pdfs[i], fs[i] = p[0], f
            L.orc_bsdf_sample(po.fp(m), po.fp(sd), po.fp(u[i]), po.fp(wi), po.fp(f), po.fp(p))
            up += wi[2] > 0
        integral = float(np.mean(pdfs.astype(np.float64)) * 2 * np.pi)
        if kind == 5:
            assert abs(integral - 1.0) < 0.02, (kind, rough, integral)
            assert up == n
        else:
            assert abs(integral - up / n) < 0.02, (kind, rough, integral, up / n)
        if kind == 5 and rough == 0.0:
            assert np.allclose(fs, np.array([0.8, 0.6, 0.4], np.float32) / np.float32(np.pi), rtol=1e-6)


def test_texture_sample_matches_reference(po):
    """Texture::sample (Imaging.h:82-108): bilinear with wrap, |u| and |v|, bit for bit
    (no transcendentals) on 1x1, 5x3 and 16x9 textures, uv far outside [0, 1]."""
    d = np.load(GOLDEN / "ref_texture.npz")
    for i in range(3):
        t, uv, ref = d[f"tex{i}"], d[f"uv{i}"], d[f"out{i}"]
        out = np.empty_like(ref)
        po.lib().orc_texture_sample(po.fp(np.ascontiguousarray(t)), t.shape[1], t.shape[0],
                                    po.fp(np.ascontiguousarray(uv)), uv.shape[0], po.fp(out))
        assert np.array_equal(bits(out), bits(ref)), i


def test_tonemap_matches_reference(po):
    """Film::tonemap (Imaging.h:220-244): extended Reinhard + gamma to 8 bits, for three
    (SPP, exposure, white point) settings, including black and near-black pixels."""
    import ctypes as C
    d = np.load(GOLDEN / "ref_tonemap.npz")
    rgb = np.ascontiguousarray(d["rgb"])
    for spp, expo, lmax in d["params"]:
        out = np.empty((rgb.shape[0], 3), np.uint8)
        po.lib().orc_tonemap(po.fp(rgb), rgb.shape[0], int(spp), float(expo), float(lmax),
                             out.ctypes.data_as(C.POINTER(C.c_uint8)))
        assert np.array_equal(out, d[f"spp{int(spp)}"]), spp


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.skipif(not SCENES.exists(), reason="the reference's scene files are not on this machine")
@pytest.mark.parametrize("name", json.load(open(GOLDEN / "ref_loader.json")).keys())
def test_loader_matches_reference(pg, name):
    """pg_scene_load of a complete reference scene directory gives the reference's
    triangles (positions, normals, uv), material indices, material records, texture
    assignment and camera bit for bit; its PNG textures decode to the texels the
    reference's Texture::load produces, its JPEG textures to within a few 8-bit levels
    (tests/golden/ref_jpeg_samples.npz)."""
    exp = json.load(open(GOLDEN / "ref_loader.json"))[name]
    d = pg.load_scene(SCENES / name)
    assert d.n_tris == exp["n_tris"] and len(d.materials) == exp["n_materials"]
    assert np.array_equal(d.camera_array(), np.array(exp["camera"], np.float32))
    assert _sha(d.positions) == exp["positions"]
    assert _sha(d.normals) == exp["normals"]
    assert _sha(d.uv) == exp["uv"]
    assert _sha(d.material_index) == exp["material_index"]
    m = d.materials_f32().copy()
    m[:, 1:4] = 0.0  # the reference's records carry the albedo in the texture only
    assert _sha(m) == exp["materials"]
    assert d.material_texture.tolist() == exp["material_texture"]
    names = [Path(p).name for p in exp["textures"]]
    assert len(d.textures) == len(names)
    js = np.load(GOLDEN / "ref_jpeg_samples.npz")
    for i, nm in enumerate(names):
        e = exp["decoded"].get(nm)
        if e is None:  # not decodable by the reference either: the white default
            assert d.textures[i].shape == (1, 1, 3) and any(nm in w for w in d.warnings), nm
            continue
        assert list(d.textures[i].shape[:2]) == e["shape"], nm
        if "sha256" in e:
            assert _sha(d.textures[i]) == e["sha256"], nm
        else:
            # JPEG: lossy, IDCT and colour conversion are not bit-specified; the
            # decoders agree on >= 95% of samples exactly and all within 4 levels
            yx, ref8 = js[e["jpeg_samples"] + ":yx"], js[e["jpeg_samples"] + ":rgb8"].astype(int)
            got8 = np.rint(d.textures[i][yx[:, 0], yx[:, 1]] * 255).astype(int)
            diff = np.abs(got8 - ref8)
            assert diff.max() <= 4 and np.mean(diff == 0) >= 0.9, (nm, diff.max(), np.mean(diff == 0))


def test_scene_directory_round_trip(pg, tmp_path):
    """A scene written with the library's own .gem / PNG / HDR writers loads back to
    what was written: instance transforms in the reference's FP32 order, normals by
    the inverse transpose, uv, one material per instance (the unknown bsdf skipped),
    the coated glass as a layered material, textures as bytes / 255, the HDR sky to
    RGBE precision, and the camera."""
    root = make_scene_dir(pg, tmp_path / "scene")
    d = pg.load_scene(root)
    assert any("velvet" in w for w in d.warnings)
    types = d.materials["type"].astype(int).tolist()
    assert types == [0, 5, 6, 2, 3 + 8, 6, 0]
    assert d.materials["roughness"][1] == np.float32(0.6)  # Oren-Nayar sigma as written
    # 2 (floor) + 2 (wall) + 12 + 12 + 12 (boxes) + 2 (two-mesh file) + 2 (light)
    assert d.n_tris == 44
    assert d.camera["width"] == 96 and d.camera["height"] == 64 and d.camera["fov"] == np.float32(50.0)
    # floor: scale 3 about the origin, uv as written
    floor = d.positions[:2]
    assert np.array_equal(floor[0], np.array([[-3, 0, -3], [3, 0, -3], [3, 0, 3]], np.float32))
    assert np.array_equal(d.uv[0], np.array([[0, 0], [1, 0], [1, 1]], np.float32))
    # the emissive quad faces down (inverse transpose of a -1 y scale)
    light = np.where(d.material_index == 6)[0]
    assert light.size == 2 and np.allclose(d.normals[light], [0, -1, 0])
    assert np.array_equal(d.materials["emission"][6], np.array([30, 25, 20], np.float32))
    # checker texture: bytes * (1 / 255) in FP32
    chk = d.textures[d.material_texture[0]]
    assert chk.shape == (8, 8, 3)
    assert np.array_equal(bits(chk[0, 0]), bits(np.array([230, 200, 90], np.float32) * np.float32(1 / 255)))
    # HDR sky: RGBE keeps ~8 bits of mantissa
    assert d.env.shape == (32, 64, 3)
    v = (np.arange(32) + 0.5) / 32
    sky = (np.array([0.35, 0.5, 0.9]) * np.clip(np.cos(np.pi * v), 0.05, 1)[:, None]).astype(np.float32)
    # truncated 8-bit mantissas under the largest component's exponent: error <= max / 128
    assert np.all(np.abs(d.env[:, 5] - sky) <= sky.max(-1, keepdims=True) / 128)
    assert np.all(np.abs(d.env[3, 40] - np.array([400, 380, 300], np.float32)) <= 400 / 128)


def test_missing_mesh_fails_or_is_skipped(pg, tmp_path):
    root = make_scene_dir(pg, tmp_path / "scene",
                          extra_instances=({"filename": "Absent.gem", "world": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0,
                                                                                0, 0, 0, 1],
                                            "bsdf": "diffuse", "reflectance": "grey.png"},))
    with pytest.raises(pg.PgError, match="Absent.gem"):
        pg.load_scene(root)
    d = pg.load_scene(root, skip_missing=True)
    assert d.n_tris == 44 and any("Absent.gem" in w for w in d.warnings)
    with pytest.raises(pg.PgError):
        pg.load_scene(tmp_path / "nowhere")


def test_image_writers_round_trip(pg, tmp_path):
    """PNG written and read back is exact; HDR written and read back is the RGBE
    quantisation of the input (max component's mantissa truncated to 8 bits); where
    the reference driver is built, the reference's stb decoder reads both files to the
    same values."""
    from oracle import pg_oracle as po
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (13, 37, 3), dtype=np.uint8)
    pg.write_png(tmp_path / "a.png", img)
    back = pg.read_image(tmp_path / "a.png")
    assert back.shape == (13, 37, 3) and np.array_equal(back, img.astype(np.float32))
    hdr = (rng.random((9, 40, 3)) ** 4 * 1e3).astype(np.float32)
    hdr[0, 0] = 0.0
    pg.write_hdr(tmp_path / "a.hdr", hdr)
    hb = pg.read_image(tmp_path / "a.hdr")
    assert hb.shape == (9, 40, 3)
    mx = hdr.max(-1, keepdims=True)
    assert np.all(np.abs(hb - hdr) <= mx / 128 + 1e-30)
    if po.ref_available():
        import ctypes as C
        R = po.ref()
        for name, want in (("a.png", img.astype(np.float32) * np.float32(1 / 255)), ("a.hdr", hb)):
            w, h = C.c_int(), C.c_int()
            R.ref_texture_load(str(tmp_path / name).encode(), C.byref(w), C.byref(h), None)
            buf = np.empty((h.value, w.value, 3), np.float32)
            R.ref_texture_load(str(tmp_path / name).encode(), C.byref(w), C.byref(h), po.fp(buf))
            assert np.array_equal(bits(buf), bits(want)), name


def test_oracle_renders_textured_scene(pg, tmp_path):
    """The oracle renders a loaded scene with textures and the new BSDFs: finite,
    lit, and the floor's checker texture shows in the image."""
    from oracle.pg_oracle import OracleScene
    d = pg.load_scene(make_scene_dir(pg, tmp_path / "scene", width=48, height=32))
    osc = OracleScene(d)
    film, st = osc.render(4, seed=2, max_depth=8, accum_f64=True, env_mode=1)
    assert np.isfinite(film).all() and film.mean() > 0
    flat = pg.SceneData(d.positions, d.normals, d.material_index, d.materials, d.env, d.camera)
    film2, _ = OracleScene(flat).render(4, seed=2, max_depth=8, accum_f64=True, env_mode=1)
    assert not np.allclose(film, film2)
