"""Host-side mirror of the reference's Scene / Camera / RayTracer interfaces.

The classes keep the reference's method names (Scene.h:74-176, Renderer.h:236-751)
and flatten into the C ABI of ``libpgtrace.so``; every render runs on the GPU.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PgCameraDesc, PgMaterial, PgRenderParams, PgStats, check, load

BSDF_DIFFUSE, BSDF_MIRROR, BSDF_CONDUCTOR, BSDF_GLASS, BSDF_DIELECTRIC, BSDF_ORENNAYAR, BSDF_PLASTIC = range(7)
BSDF_LAYERED = 8  # added to a base type: LayeredBSDF over that base (Materials.h:855-897)

MATERIAL_DTYPE = np.dtype([("type", "f4"), ("albedo", "f4", 3), ("emission", "f4", 3), ("eta", "f4", 3),
                           ("k", "f4", 3), ("int_ior", "f4"), ("ext_ior", "f4"), ("roughness", "f4")])


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


@dataclass
class SceneData:
    """Triangle soup + materials + optional environment map + camera."""

    positions: np.ndarray          # (n, 3, 3) float32: v0, v1, v2
    normals: np.ndarray            # (n, 3, 3) float32: per-vertex shading normals
    material_index: np.ndarray     # (n,) uint32
    materials: np.ndarray          # (m,) MATERIAL_DTYPE
    env: np.ndarray | None         # (h, w, 3) float32 lat-long, y-up
    camera: dict = field(default_factory=dict)  # from, to, up, fov, width, height
    # albedo textures (scene files; None for procedural scenes)
    uv: np.ndarray | None = None                 # (n, 3, 2) float32 per-vertex (u, v)
    material_texture: np.ndarray | None = None   # (m,) int32, -1 = constant albedo
    textures: list = field(default_factory=list)  # (h, w, 3) float32 RGB each
    warnings: list = field(default_factory=list)

    @property
    def has_textures(self) -> bool:
        return self.material_texture is not None and len(self.textures) > 0

    def texture_arrays(self):
        """uv (n*6), material_texture, concatenated RGB texels, offsets, widths, heights."""
        uv = np.ascontiguousarray(self.uv if self.uv is not None else np.zeros((self.n_tris, 3, 2)), np.float32)
        mt = np.ascontiguousarray(self.material_texture if self.material_texture is not None
                                  else np.full(len(self.materials), -1), np.int32)
        texels = np.ascontiguousarray(np.concatenate([t.reshape(-1, 3) for t in self.textures])
                                      if self.textures else np.zeros((0, 3)), np.float32)
        sizes = [t.shape[0] * t.shape[1] for t in self.textures]
        off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]) if sizes else np.zeros(0), np.uint64)
        tw = np.array([t.shape[1] for t in self.textures], np.int32)
        th = np.array([t.shape[0] for t in self.textures], np.int32)
        return uv, mt, texels, off, tw, th

    @property
    def n_tris(self) -> int:
        return int(self.positions.shape[0])

    @property
    def width(self) -> int:
        return int(self.camera["width"])

    @property
    def height(self) -> int:
        return int(self.camera["height"])

    def camera_array(self) -> np.ndarray:
        c = self.camera
        return np.array([*c["from"], *c["to"], *c["up"], c["fov"], c["width"], c["height"]], dtype=np.float32)

    def materials_f32(self) -> np.ndarray:
        return np.ascontiguousarray(self.materials).view(np.float32).reshape(-1, 16)


def generate_scene(config: int, seed: int = 0) -> SceneData:
    """Procedural scene of a benchmark configuration (1 Cornell box, 2 open interior,
    3 interior with a window, 5 batched-view family, 100/101 small test interiors)."""
    lib = load()
    h = C.c_void_p()
    check(lib.pg_scene_generate(int(config), int(seed), C.byref(h)))
    try:
        nt, nm = C.c_uint32(), C.c_uint32()
        ew, eh = C.c_int32(), C.c_int32()
        check(lib.pg_scene_sizes(h, C.byref(nt), C.byref(nm), C.byref(ew), C.byref(eh)))
        pos = np.empty((nt.value, 3, 3), np.float32)
        nrm = np.empty((nt.value, 3, 3), np.float32)
        mi = np.empty(nt.value, np.uint32)
        mats = np.zeros(nm.value, MATERIAL_DTYPE)
        env = np.empty((eh.value, ew.value, 3), np.float32) if ew.value > 0 else None
        cam = PgCameraDesc()
        check(lib.pg_scene_copy(h, _fp(pos), _fp(nrm), _up(mi), mats.ctypes.data_as(C.POINTER(PgMaterial)),
                                _fp(env) if env is not None else None, C.byref(cam)))
    finally:
        lib.pg_scene_free(h)
    camera = {"from": list(cam.from_), "to": list(cam.to), "up": list(cam.up), "fov": cam.fov_deg,
              "width": int(cam.width), "height": int(cam.height)}
    return SceneData(pos, nrm, mi, mats, env, camera)


def load_scene(path, skip_missing: bool = False) -> SceneData:
    """loadScene (SceneLoader.h:185-269): a reference scene directory (scene.json,
    .gem meshes, PNG / HDR textures) as a SceneData with its albedo textures.
    skip_missing drops instances whose mesh file is absent instead of failing."""
    lib = load()
    h = C.c_void_p()
    code = lib.pg_scene_load(str(path).encode(), _lib.PG_LOAD_SKIP_MISSING if skip_missing else 0, C.byref(h))
    if code != _lib.PG_OK:
        raise _lib.PgError(code, (lib.pg_scene_load_error() or b"").decode())
    try:
        nt, nm = C.c_uint32(), C.c_uint32()
        ew, eh = C.c_int32(), C.c_int32()
        check(lib.pg_scene_sizes(h, C.byref(nt), C.byref(nm), C.byref(ew), C.byref(eh)))
        ntex, nwarn = C.c_uint32(), C.c_uint32()
        ntexel = C.c_uint64()
        check(lib.pg_scene_texture_sizes(h, C.byref(ntex), C.byref(ntexel), C.byref(nwarn)))
        pos = np.empty((nt.value, 3, 3), np.float32)
        nrm = np.empty((nt.value, 3, 3), np.float32)
        mi = np.empty(nt.value, np.uint32)
        mats = np.zeros(nm.value, MATERIAL_DTYPE)
        env = np.empty((eh.value, ew.value, 3), np.float32) if ew.value > 0 else None
        cam = PgCameraDesc()
        check(lib.pg_scene_copy(h, _fp(pos), _fp(nrm), _up(mi), mats.ctypes.data_as(C.POINTER(PgMaterial)),
                                _fp(env) if env is not None else None, C.byref(cam)))
        uv = np.empty((nt.value, 3, 2), np.float32)
        mt = np.empty(nm.value, np.int32)
        texels = np.empty((ntexel.value, 3), np.float32)
        off = np.empty(ntex.value, np.uint64)
        tw = np.empty(ntex.value, np.int32)
        th = np.empty(ntex.value, np.int32)
        check(lib.pg_scene_copy_textures(h, _fp(uv), mt.ctypes.data_as(C.POINTER(C.c_int32)), _fp(texels),
                                         off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         tw.ctypes.data_as(C.POINTER(C.c_int32)),
                                         th.ctypes.data_as(C.POINTER(C.c_int32))))
        warnings = [lib.pg_scene_warning(h, i).decode() for i in range(nwarn.value)]
    finally:
        lib.pg_scene_free(h)
    textures = [texels[int(o): int(o) + int(w) * int(hh)].reshape(int(hh), int(w), 3).copy()
                for o, w, hh in zip(off, tw, th)]
    camera = {"from": list(cam.from_), "to": list(cam.to), "up": list(cam.up), "fov": cam.fov_deg,
              "width": int(cam.width), "height": int(cam.height)}
    return SceneData(pos, nrm, mi, mats, env, camera, uv=uv, material_texture=mt, textures=textures,
                     warnings=warnings)


def read_image(path) -> np.ndarray:
    """Decode a PNG (8-bit samples as 0..255 floats) or Radiance HDR file as the
    scene loader does: (h, w, channels) float32."""
    lib = load()
    w, h, c = C.c_int32(), C.c_int32(), C.c_int32()
    p = str(path).encode()
    code = lib.pg_image_read(p, C.byref(w), C.byref(h), C.byref(c), None)
    if code != _lib.PG_OK:
        raise _lib.PgError(code, (lib.pg_scene_load_error() or b"").decode())
    out = np.empty((h.value, w.value, c.value), np.float32)
    check(lib.pg_image_read(p, C.byref(w), C.byref(h), C.byref(c), _fp(out)))
    return out


def write_png(path, rgb8: np.ndarray) -> None:
    """8-bit RGB PNG (RayTracer::savePNG, Renderer.h:751)."""
    a = np.ascontiguousarray(rgb8, np.uint8)
    check(load().pg_image_write_png(str(path).encode(), a.ctypes.data_as(C.POINTER(C.c_uint8)), a.shape[1],
                                    a.shape[0]))


def write_hdr(path, rgb: np.ndarray) -> None:
    """Radiance RGBE image of an RGB float image (Film::save, Imaging.h:263-272)."""
    a = np.ascontiguousarray(rgb, np.float32)
    check(load().pg_image_write_hdr(str(path).encode(), _fp(a), a.shape[1], a.shape[0]))


# Contexts and batches still open when the interpreter exits are closed by an atexit
# hook, batches first: module teardown would otherwise release them in no fixed order,
# possibly after the HIP runtime has begun to shut down.
_open_devices: "weakref.WeakSet" = weakref.WeakSet()
_open_batches: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_all():
    for b in list(_open_batches):
        b.close()
    for d in list(_open_devices):
        d.close()


class Device:
    """A device context (pg_ctx) bound to one GPU."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.ctx = C.c_void_p()
        check(self.lib.pg_create(C.byref(self.ctx), int(device)))
        self.current_scene = None  # the Scene whose upload the context holds
        _open_devices.add(self)

    def close(self):
        if self.ctx:
            self.lib.pg_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code: int):
        check(code, self.ctx)

    # ---------------------------------------------------------------- guiding cache
    def records(self) -> np.ndarray:
        """PathVertex records (n, 12) float32: pos, normal, wi, Li."""
        n = C.c_uint64()
        self._chk(self.lib.pg_records_count(self.ctx, C.byref(n)))
        out = np.empty((n.value, 12), np.float32)
        got = C.c_uint64()
        self._chk(self.lib.pg_records_read(self.ctx, out.ctypes.data_as(C.c_void_p), n.value, C.byref(got)))
        return out[: got.value]

    def upload_records(self, rec: np.ndarray):
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 12)
        self._chk(self.lib.pg_records_upload(self.ctx, rec.ctypes.data_as(C.c_void_p), rec.shape[0]))

    def clear_records(self):
        self._chk(self.lib.pg_records_clear(self.ctx))

    def cache_build(self, cell_size: int = 4096, qt_nodes: int = 1024):
        self._chk(self.lib.pg_cache_build(self.ctx, cell_size, qt_nodes))
        return self.cache_info()

    def cache_info(self) -> dict:
        n = C.c_uint64()
        nodes, depth, cells = C.c_uint32(), C.c_uint32(), C.c_uint32()
        ms = C.c_double()
        self._chk(self.lib.pg_cache_info(self.ctx, C.byref(n), C.byref(nodes), C.byref(depth), C.byref(cells),
                                         C.byref(ms)))
        return {"records": n.value, "nodes": nodes.value, "depth": depth.value, "cells": cells.value,
                "build_ms": ms.value}

    def cache_export(self, qt_nodes: int = 1024):
        info = self.cache_info()
        nodes = np.empty((info["nodes"], 8), np.float32)
        pos = np.empty((info["records"], 4), np.float32)
        qt = np.empty((max(info["cells"], 1) * qt_nodes, 8), np.float32)
        self._chk(self.lib.pg_cache_export(self.ctx, _fp(nodes), _fp(pos), _fp(qt)))
        return nodes, pos, qt

    def knn(self, q: np.ndarray, k: int = 16):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        idx = np.empty((q.shape[0], k), np.uint32)
        d2 = np.empty((q.shape[0], k), np.float32)
        self._chk(self.lib.pg_knn(self.ctx, _fp(q), q.shape[0], k, _up(idx), _fp(d2)))
        return idx, d2

    def guide_eval(self, pos: np.ndarray, u3: np.ndarray, wi_eval: np.ndarray):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        u3 = np.ascontiguousarray(u3, np.float32).reshape(-1, 3)
        we = np.ascontiguousarray(wi_eval, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        wi = np.empty((n, 3), np.float32)
        ps, pe = np.empty(n, np.float32), np.empty(n, np.float32)
        self._chk(self.lib.pg_guide_qt_eval(self.ctx, _fp(pos), _fp(u3), _fp(we), n, _fp(wi), _fp(ps), _fp(pe)))
        return wi, ps, pe

    def guide_pss_eval(self, material: int, pos, nrm, wo, u3, wi_eval):
        """Primary-sample-space guide at shading points of one material
        (pg_guide_pss_eval): sampled directions, their pdf, and the pdf of wi_eval."""
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1, 3) for a in (pos, nrm, wo, u3, wi_eval)]
        n = arrs[0].shape[0]
        if any(a.shape[0] != n for a in arrs):
            raise ValueError("pos, nrm, wo, u3 and wi_eval need one row per point")
        wi = np.empty((n, 3), np.float32)
        ps, pe = np.empty(n, np.float32), np.empty(n, np.float32)
        self._chk(self.lib.pg_guide_pss_eval(self.ctx, int(material), *(_fp(a) for a in arrs), n, _fp(wi), _fp(ps),
                                             _fp(pe)))
        return wi, ps, pe

    # ---------------------------------------------------------------- learned guide
    def mlp_config(self, **kw) -> _lib.PgMlpConfig:
        cfg = self.lib.pg_mlp_config_default()
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise KeyError(k)
            setattr(cfg, k, v)
        return cfg

    def mlp_init(self, **kw):
        cfg = self.mlp_config(**kw)
        self._chk(self.lib.pg_mlp_init(self.ctx, C.byref(cfg)))
        return self.mlp_info()

    def mlp_train(self, steps: int) -> np.ndarray:
        losses = np.empty(max(steps, 1), np.float32)
        self._chk(self.lib.pg_mlp_train(self.ctx, int(steps), _fp(losses)))
        return losses[:steps]

    def mlp_info(self) -> dict:
        i = _lib.PgMlpInfo()
        self._chk(self.lib.pg_mlp_info_read(self.ctx, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_ if k != "reserved"}

    def mlp_params(self):
        info = self.mlp_info()
        p = np.empty(info["n_params"], np.float32)
        g = np.empty(info["n_grid_params"], np.float32)
        self._chk(self.lib.pg_mlp_get_params(self.ctx, _fp(p), _fp(g)))
        return p, g

    def mlp_set_params(self, params=None, grid=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        self._chk(self.lib.pg_mlp_set_params(self.ctx, _fp(p) if p is not None else None,
                                             _fp(g) if g is not None else None))

    def mlp_levels(self):
        res = np.empty(16, np.uint32)
        dense = np.empty(16, np.uint32)
        frame = np.empty(4, np.float32)
        self._chk(self.lib.pg_mlp_levels(self.ctx, _up(res), _up(dense), _fp(frame)))
        return res, dense.astype(bool), frame

    def mlp_eval(self, pos, nrm, wo, wi_eval=None, u3=None):
        """Mixture (n, 13, 4), pdf of wi_eval (n,), sampled directions (n, 4: xyz, pdf)."""
        f3 = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1, 3)
        pos, nrm, wo = f3(pos), f3(nrm), f3(wo)
        n = pos.shape[0]
        mix = np.empty((n, 13, 4), np.float32)
        we = f3(wi_eval) if wi_eval is not None else None
        pe = np.empty(n, np.float32) if wi_eval is not None else None
        uu = f3(u3) if u3 is not None else None
        wo_out = np.empty((n, 4), np.float32) if u3 is not None else None
        opt = lambda a: _fp(a) if a is not None else None
        self._chk(self.lib.pg_mlp_eval(self.ctx, n, _fp(pos), _fp(nrm), _fp(wo), _fp(mix), opt(we), opt(pe),
                                       opt(uu), opt(wo_out)))
        return mix, pe, wo_out

    def mlp_gradient(self, samples, want_grid: bool = True):
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 16)
        info = self.mlp_info()
        loss = np.zeros(1, np.float32)
        gp = np.empty(info["n_params"], np.float32)
        gg = np.empty(info["n_grid_params"], np.float32) if want_grid else None
        self._chk(self.lib.pg_mlp_gradient(self.ctx, s.shape[0], _fp(s), _fp(loss), _fp(gp),
                                           _fp(gg) if gg is not None else None))
        return float(loss[0]), gp, gg

    def bsdf_query(self, material: np.ndarray, sd: np.ndarray, u: np.ndarray, wi_eval: np.ndarray) -> np.ndarray:
        """BSDF sample / evaluate / PDF / invert and flags of one material (a 16-float
        record) on the device with given random numbers: (n, 15) float32."""
        m = PgMaterial.from_buffer_copy(np.ascontiguousarray(material, np.float32).tobytes())
        sd = np.ascontiguousarray(sd, np.float32).reshape(-1, 12)
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 3)
        we = np.ascontiguousarray(wi_eval, np.float32).reshape(-1, 3)
        out = np.empty((sd.shape[0], 15), np.float32)
        self._chk(self.lib.pg_bsdf_query(self.ctx, C.byref(m), sd.shape[0], _fp(sd), _fp(u), _fp(we), _fp(out)))
        return out

    def info(self) -> dict:
        name = C.create_string_buffer(64)
        cu = C.c_int32()
        peak = C.c_double()
        self._chk(self.lib.pg_device_info(self.ctx, name, 64, C.byref(cu), C.byref(peak)))
        return {"name": name.value.decode(), "cu_count": cu.value, "peak_hbm_gbs": peak.value}


class Scene:
    """Scene::init + Scene::build (Scene.h:84-137) on the device; traverse / visible as
    batched queries (Scene.h:103-146). A device context holds one uploaded scene: once
    another Scene is created on the same Device, this one and its RayTracers raise
    instead of rendering the other scene."""

    def __init__(self, device: Device, data: SceneData):
        self.device = device
        self.data = data
        lib = device.lib
        device.current_scene = None
        pos = np.ascontiguousarray(data.positions, np.float32)
        nrm = np.ascontiguousarray(data.normals, np.float32)
        mi = np.ascontiguousarray(data.material_index, np.uint32)
        mats = np.ascontiguousarray(data.materials)
        env = None if data.env is None else np.ascontiguousarray(data.env, np.float32)
        c = data.camera
        cam = PgCameraDesc((C.c_float * 3)(*c["from"]), (C.c_float * 3)(*c["to"]), (C.c_float * 3)(*c["up"]),
                           c["fov"], float(c["width"]), float(c["height"]))
        device._chk(lib.pg_scene_upload(device.ctx, _fp(pos), _fp(nrm), _up(mi), pos.shape[0],
                                        mats.ctypes.data_as(C.POINTER(PgMaterial)), mats.shape[0],
                                        _fp(env) if env is not None else None,
                                        0 if env is None else env.shape[1], 0 if env is None else env.shape[0],
                                        C.byref(cam)))
        if data.has_textures:
            uv, mt, texels, off, tw, th = data.texture_arrays()
            device._chk(lib.pg_scene_textures(device.ctx, _fp(uv), mt.ctypes.data_as(C.POINTER(C.c_int32)),
                                              mt.shape[0], _fp(texels), off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              tw.ctypes.data_as(C.POINTER(C.c_int32)),
                                              th.ctypes.data_as(C.POINTER(C.c_int32)), len(data.textures)))
        device.current_scene = self

    def check_current(self):
        """Raise unless the device context holds this scene's upload."""
        if self.device.current_scene is not self:
            raise RuntimeError("another Scene was uploaded to this Device since: the context renders and "
                               "traces that one (create this Scene again to use it)")

    def bvh_info(self) -> dict:
        self.check_current()
        n, d, ms = C.c_uint32(), C.c_uint32(), C.c_double()
        self.device._chk(self.device.lib.pg_bvh_info(self.device.ctx, C.byref(n), C.byref(d), C.byref(ms)))
        return {"nodes": n.value, "max_depth": d.value, "build_ms": ms.value}

    def bvh_export(self):
        self.check_current()
        info = self.bvh_info()
        nodes = np.empty((info["nodes"], 8), np.float32)
        idx = np.empty(self.data.n_tris, np.int32)
        self.device._chk(self.device.lib.pg_bvh_export(self.device.ctx, _fp(nodes),
                                                        idx.ctypes.data_as(C.POINTER(C.c_int32))))
        return nodes, idx

    def traverse(self, rays: np.ndarray):
        """Closest hits for rays (n, 6) = origin, direction. Returns id, t, alpha, beta, gamma."""
        self.check_current()
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        hits = np.empty(n, np.dtype([("id", "u4"), ("t", "f4"), ("alpha", "f4"), ("beta", "f4"), ("gamma", "f4")]))
        self.device._chk(self.device.lib.pg_trace_closest(self.device.ctx, _fp(rays), n,
                                                           hits.ctypes.data_as(C.c_void_p)))
        return hits["id"].copy(), hits["t"].copy(), hits["alpha"].copy(), hits["beta"].copy(), hits["gamma"].copy()

    def visible_rays(self, rays: np.ndarray, max_t: np.ndarray) -> np.ndarray:
        self.check_current()
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        max_t = np.ascontiguousarray(max_t, np.float32)
        out = np.empty(rays.shape[0], np.uint8)
        self.device._chk(self.device.lib.pg_trace_shadow(self.device.ctx, _fp(rays), _fp(max_t), rays.shape[0],
                                                          out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def generate_rays(self, pxy: np.ndarray) -> np.ndarray:
        self.check_current()
        pxy = np.ascontiguousarray(pxy, np.float32).reshape(-1, 2)
        out = np.empty((pxy.shape[0], 6), np.float32)
        self.device._chk(self.device.lib.pg_generate_rays(self.device.ctx, _fp(pxy), pxy.shape[0], _fp(out)))
        return out


class RayTracer:
    """RayTracer (Renderer.h:236-751): init / clear / render / getSPP / film read-back."""

    def __init__(self, scene: Scene, **params):
        self.scene = scene
        self.device = scene.device
        self.params = self.device.lib.pg_render_params_default()
        for k, v in params.items():
            setattr(self.params, k, v)

    def clear(self):
        self.scene.check_current()
        self.device._chk(self.device.lib.pg_film_clear(self.device.ctx))

    def render(self, n_spp: int = 1):
        self.scene.check_current()
        self.device._chk(self.device.lib.pg_render_spp(self.device.ctx, C.byref(self.params), int(n_spp)))

    def render_online(self, n_spp: int, steps: int = 1, window: int = 1, train_every: int = 1,
                      warmup_passes: int = 0, warmup_steps: int = 0, record_samples: int = 1):
        """Learned-guide passes with online training (pg_mlp_render_online)."""
        self.scene.check_current()
        sch = _lib.PgMlpSchedule(train_every, steps, window, warmup_passes, warmup_steps, record_samples)
        self.device._chk(self.device.lib.pg_mlp_render_online(self.device.ctx, C.byref(self.params), int(n_spp),
                                                              C.byref(sch)))

    def tune(self) -> float:
        """Pick the fastest launch shape for these parameters (pg_render_tune); clears
        the film. Returns the tuned milliseconds per SPP."""
        self.scene.check_current()
        tuned = PgRenderParams()
        ms = C.c_double()
        self.device._chk(self.device.lib.pg_render_tune(self.device.ctx, C.byref(self.params), C.byref(tuned),
                                                        C.byref(ms)))
        self.params = tuned
        return ms.value

    def getSPP(self) -> int:  # noqa: N802 - reference name
        _, spp = self.film()
        return spp

    def film(self):
        """(H, W, 3) float32 sum over samples and the SPP count."""
        self.scene.check_current()
        W, H = self.scene.data.width, self.scene.data.height
        buf = np.empty((H, W, 3), np.float32)
        spp = C.c_uint32()
        self.device._chk(self.device.lib.pg_film_read(self.device.ctx, _fp(buf), C.byref(spp)))
        return buf, spp.value

    def image(self) -> np.ndarray:
        """Film divided by SPP (Film::save, Imaging.h:263-272)."""
        buf, spp = self.film()
        return buf * (1.0 / spp if spp > 0 else 1.0)

    def tonemap(self, exposure: float = 1.0, lmax_white: float = 10.0) -> np.ndarray:
        """Film::tonemap (Imaging.h:220-244) of every pixel on the device: (H, W, 3) uint8."""
        W, H = self.scene.data.width, self.scene.data.height
        out = np.empty((H, W, 3), np.uint8)
        self.device._chk(self.device.lib.pg_film_tonemap(self.device.ctx, float(exposure), float(lmax_white),
                                                         out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def savePNG(self, filename) -> None:  # noqa: N802 - reference name (Renderer.h:751)
        write_png(filename, self.tonemap())

    def saveHDR(self, filename) -> None:  # noqa: N802 - reference name (Renderer.h:750)
        write_hdr(filename, self.image())

    def stats(self) -> dict:
        s = PgStats()
        self.device._chk(self.device.lib.pg_stats_read(self.device.ctx, C.byref(s)))
        return {f: getattr(s, f) for f, _ in PgStats._fields_}

    def reset_stats(self):
        self.device._chk(self.device.lib.pg_stats_reset(self.device.ctx))

    def set_timing(self, on: bool):
        """Per-launch kernel timing (HIP events around every render launch)."""
        self.device._chk(self.device.lib.pg_set_timing(self.device.ctx, 1 if on else 0))

    def set_counting(self, on: bool):
        self.device._chk(self.device.lib.pg_set_counting(self.device.ctx, 1 if on else 0))


def _camera_desc(c: dict) -> PgCameraDesc:
    return PgCameraDesc((C.c_float * 3)(*c["from"]), (C.c_float * 3)(*c["to"]), (C.c_float * 3)(*c["up"]),
                        c["fov"], float(c["width"]), float(c["height"]))


class Batch:
    """Many views of one or more scenes rendered together (pg_batch_*): one hipGraph
    launch per SPP, view-aligned chunks, autotuned launch configuration."""

    def __init__(self, device: Device):
        self.device = device
        self.lib = device.lib
        self.h = C.c_void_p()
        device._chk(self.lib.pg_batch_create(device.ctx, C.byref(self.h)))
        self.views = []  # (scene id, width, height)
        self._keep = []
        _open_batches.add(self)

    def close(self):
        if self.h:
            self.lib.pg_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, data: SceneData) -> int:
        pos = np.ascontiguousarray(data.positions, np.float32)
        nrm = np.ascontiguousarray(data.normals, np.float32)
        mi = np.ascontiguousarray(data.material_index, np.uint32)
        mats = np.ascontiguousarray(data.materials)
        env = None if data.env is None else np.ascontiguousarray(data.env, np.float32)
        sid = C.c_uint32()
        self.device._chk(self.lib.pg_batch_add_scene(self.h, _fp(pos), _fp(nrm), _up(mi), pos.shape[0],
                                                     mats.ctypes.data_as(C.POINTER(PgMaterial)), mats.shape[0],
                                                     _fp(env) if env is not None else None,
                                                     0 if env is None else env.shape[1],
                                                     0 if env is None else env.shape[0], C.byref(sid)))
        if data.has_textures:
            uv, mt, texels, off, tw, th = data.texture_arrays()
            self.device._chk(self.lib.pg_batch_scene_textures(
                self.h, sid.value, _fp(uv), mt.ctypes.data_as(C.POINTER(C.c_int32)), mt.shape[0], _fp(texels),
                off.ctypes.data_as(C.POINTER(C.c_uint64)), tw.ctypes.data_as(C.POINTER(C.c_int32)),
                th.ctypes.data_as(C.POINTER(C.c_int32)), len(data.textures)))
        return sid.value

    def add_view(self, scene_id: int, camera: dict) -> int:
        vid = C.c_uint32()
        cam = _camera_desc(camera)
        self.device._chk(self.lib.pg_batch_add_view(self.h, int(scene_id), C.byref(cam), C.byref(vid)))
        self.views.append((scene_id, int(camera["width"]), int(camera["height"])))
        return vid.value

    def set_guide(self, scene_id: int, params: np.ndarray, grid: np.ndarray, **cfg):
        """Learned guide of one scene (pg_batch_set_guide): the network parameters and
        hash grid as Device.mlp_params() returns them, and the configuration keywords
        they were trained with (Device.mlp_init's)."""
        c = self.device.mlp_config(**cfg)
        p = np.ascontiguousarray(params, np.float32)
        g = np.ascontiguousarray(grid, np.float32)
        self.device._chk(self.lib.pg_batch_set_guide(self.h, int(scene_id), C.byref(c), _fp(p), _fp(g)))

    def configure(self, **kw):
        cfg = _lib.PgBatchConfig(0, 0, 0, 0, 1, 0)
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.device._chk(self.lib.pg_batch_configure(self.h, C.byref(cfg)))

    def _params(self, kw) -> PgRenderParams:
        p = self.lib.pg_render_params_default()
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def autotune(self, **params):
        p = self._params(params)
        best = _lib.PgBatchConfig()
        ms = C.c_double()
        self.device._chk(self.lib.pg_batch_autotune(self.h, C.byref(p), C.byref(best), C.byref(ms)))
        return {k: getattr(best, k) for k, _ in best._fields_ if k != "reserved"}, ms.value

    def clear(self):
        self.device._chk(self.lib.pg_batch_clear(self.h))

    def render(self, n_spp: int = 1, **params):
        p = self._params(params)
        self.device._chk(self.lib.pg_batch_render(self.h, C.byref(p), int(n_spp)))

    def film(self, view: int):
        _, w, h = self.views[view]
        buf = np.empty((h, w, 3), np.float32)
        spp = C.c_uint32()
        self.device._chk(self.lib.pg_batch_film_read(self.h, int(view), _fp(buf), C.byref(spp)))
        return buf, spp.value

    def image(self, view: int) -> np.ndarray:
        buf, spp = self.film(view)
        return buf * (1.0 / spp if spp > 0 else 1.0)

    def stats(self) -> dict:
        s = _lib.PgBatchStats()
        self.device._chk(self.lib.pg_batch_stats_read(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


def metrics(img: np.ndarray, ref: np.ndarray) -> dict:
    """Metrics::calculateMSE / RMSE / RRMSE (Metrics.h:8-69, FP32, bit-exact) and
    relMSE (FP64) of an (H, W, 3) image against a reference, via pg_metrics."""
    a = np.ascontiguousarray(img, np.float32)
    b = np.ascontiguousarray(ref, np.float32)
    h, w = (a.shape[0], a.shape[1]) if a.ndim == 3 else (1, a.size // 3)
    out = _lib.PgImageMetrics()
    code = load().pg_metrics(_fp(a), _fp(b), a.size, b.size, w, h, C.byref(out))
    res = {"mse": out.mse, "rmse": out.rmse, "rrmse": out.rrmse, "relmse": out.relmse}
    if code != _lib.PG_OK:
        raise _lib.PgError(code, "image sizes differ")
    return res


def relmse(img: np.ndarray, ref: np.ndarray, eps: float = 1e-2) -> float:
    """Mean over pixels and RGB of (I - R)^2 / (R^2 + eps), in FP64."""
    i = img.astype(np.float64)
    r = ref.astype(np.float64)
    return float(np.mean((i - r) ** 2 / (r * r + eps)))
