"""ctypes wrapper of the CPU oracle (oracle/pg_oracle.cpp) and, where it was built,
of the reference golden-vector driver (oracle/_ref/libpgref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "build" / "libpgoracle.so"
REF_SO = HERE / "_ref" / "libpgref.so"

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)


def fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def up(a):
    return None if a is None else a.ctypes.data_as(_U32)


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("spp_offset", C.c_uint32), ("seed", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("integrator", C.c_int32), ("sampler", C.c_int32),
                ("accum_f64", C.c_int32), ("n_threads", C.c_int32), ("record", C.c_int32),
                ("env_mode", C.c_int32), ("env_visibility", C.c_int32), ("crop", C.c_int32 * 4)]


class RenderStats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("seconds", C.c_double), ("n_records", C.c_uint64)]


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference driver when the reference is present)."""
    cmd = ["make", "-s", "-C", str(HERE), "all"]
    if force:
        cmd.insert(2, "-B")
    subprocess.run(cmd, check=True)
    if Path("/root/reference/PathGuidingDissertation/Geometry.h").exists():
        cmd = ["make", "-s", "-C", str(HERE), "ref"]
        if force:
            cmd.insert(2, "-B")
        subprocess.run(cmd, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build()
        L = C.CDLL(str(ORACLE_SO))
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [_F, _F, _U32, C.c_uint32, _F, C.c_uint32, _F, C.c_int32, C.c_int32, _F]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_bvh_node_count.restype = C.c_uint32
        L.orc_bvh_node_count.argtypes = [C.c_void_p]
        L.orc_bvh_max_depth.restype = C.c_int32
        L.orc_bvh_max_depth.argtypes = [C.c_void_p]
        L.orc_bvh_export.argtypes = [C.c_void_p, _F, _I32]
        L.orc_camera_export.argtypes = [C.c_void_p, _F, _F, _F]
        L.orc_generate_rays.argtypes = [C.c_void_p, _F, C.c_uint32, _F]
        L.orc_trace_closest.argtypes = [C.c_void_p, _F, C.c_uint32, _U32, _F, _F, _F, _F, _U32, _U32, C.c_int32]
        L.orc_trace_shadow.argtypes = [C.c_void_p, _F, _F, C.c_uint32, _U8, _U32, _U32, C.c_int32]
        L.orc_trace_brute.argtypes = [C.c_void_p, _F, C.c_uint32, _U32, _F]
        L.orc_render.restype = C.c_int
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(RenderParams), _D, C.POINTER(RenderStats)]
        L.orc_records_count.restype = C.c_uint64
        L.orc_records_count.argtypes = [C.c_void_p]
        L.orc_records_export.argtypes = [C.c_void_p, _F]
        L.orc_rng.restype = C.c_float
        L.orc_rng.argtypes = [C.c_uint32] * 4
        L.orc_bsdf_sample.argtypes = [_F, _F, _F, _F, _F, _F]
        L.orc_bsdf_eval.argtypes = [_F, _F, _F, _F, _F]
        L.orc_bsdf_invert.argtypes = [_F, _F, _F, _F]
        L.orc_fresnel_dielectric.restype = C.c_float
        L.orc_fresnel_dielectric.argtypes = [C.c_float] * 3
        L.orc_fresnel_conductor.argtypes = [C.c_float, _F, _F, _F]
        L.orc_ggx_d.restype = C.c_float
        L.orc_ggx_d.argtypes = [_F, C.c_float]
        L.orc_ggx_g.restype = C.c_float
        L.orc_ggx_g.argtypes = [_F, _F, C.c_float]
        L.orc_env_sample.argtypes = [C.c_void_p, C.c_float, C.c_float, _F, _F, _F]
        L.orc_env_eval.argtypes = [C.c_void_p, _F, _F, _F]
        L.orc_env_sample_alias.argtypes = [C.c_void_p, _F, _F, _F, _F]
        L.orc_env_tables.argtypes = [C.c_void_p, _F, _U32, _F]
        L.orc_metrics.argtypes = [_F, _F, C.c_uint64, C.c_uint64, C.c_int, C.c_int, _F]
        L.orc_env_marginal_integral.restype = C.c_float
        L.orc_env_marginal_integral.argtypes = [C.c_void_p]
        L.orc_scene_bounds.argtypes = [C.c_void_p, _F, _F]
        L.orc_point_bvh_stats.restype = C.c_int32
        L.orc_point_bvh_stats.argtypes = [_F, C.c_uint32, _I32]
        L.orc_knn_brute.argtypes = [_F, C.c_uint32, _F, C.c_uint32, C.c_uint32, _U32, _F, C.c_int32]
        L.orc_scene_set_textures.argtypes = [C.c_void_p, _F, _I32, C.c_uint32, _F, C.POINTER(C.c_uint64), _I32,
                                             _I32, C.c_uint32]
        L.orc_texture_sample.argtypes = [_F, C.c_int32, C.c_int32, _F, C.c_uint32, _F]
        L.orc_bsdf_flags.argtypes = [_F, _I32, _I32]
        L.orc_tonemap.argtypes = [_F, C.c_uint64, C.c_int32, C.c_float, C.c_float, _U8]
        L.orc_lights.restype = C.c_int32
        L.orc_lights.argtypes = [C.c_void_p, _I32, _I32]
        L.orc_area_light.argtypes = [C.c_void_p, C.c_int32, _F, _F, _F]
        L.orc_sample_light.argtypes = [C.c_void_p, _F, C.c_uint32, _I32, _F]
        L.orc_shading_data.argtypes = [C.c_void_p, _F, C.c_uint32, _F]
        L.orc_visible.argtypes = [C.c_void_p, _F, _F, C.c_uint32, _U8]
        L.orc_film_splat.argtypes = [C.c_int32, C.c_int32, _F, _F, C.c_uint32, _F]
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_SO.exists()


_ref = None


def ref():
    """The reference's own classes behind C entry points (built only where the
    reference sources exist)."""
    global _ref
    if _ref is None:
        R = C.CDLL(str(REF_SO))
        R.ref_scene_create.restype = C.c_void_p
        R.ref_scene_create.argtypes = [_F, _F, C.c_uint32]
        R.ref_scene_destroy.argtypes = [C.c_void_p]
        R.ref_bvh_node_count.restype = C.c_uint32
        R.ref_bvh_node_count.argtypes = [C.c_void_p]
        R.ref_bvh_export.argtypes = [C.c_void_p, _F, _I32]
        R.ref_trace_closest.argtypes = [C.c_void_p, _F, C.c_uint32, _U32, _F, _F, _F, _F]
        R.ref_trace_shadow.argtypes = [C.c_void_p, _F, _F, C.c_uint32, _U8]
        R.ref_visible.argtypes = [C.c_void_p, _F, _F, C.c_uint32, _U8]
        R.ref_camera.argtypes = [_F, _F, C.c_uint32, _F, _F, _F]
        R.ref_bsdf.argtypes = [_F, _F, _F, _F, _F]
        R.ref_fresnel_dielectric.restype = C.c_float
        R.ref_fresnel_dielectric.argtypes = [C.c_float] * 3
        R.ref_fresnel_conductor.argtypes = [C.c_float, _F, _F, _F]
        R.ref_ggx_d.restype = C.c_float
        R.ref_ggx_d.argtypes = [_F, C.c_float]
        R.ref_ggx_g.restype = C.c_float
        R.ref_ggx_g.argtypes = [_F, _F, C.c_float]
        R.ref_env_create.restype = C.c_void_p
        R.ref_env_create.argtypes = [_F, C.c_int, C.c_int]
        R.ref_env_destroy.argtypes = [C.c_void_p]
        R.ref_env_sample.argtypes = [C.c_void_p, C.c_float, C.c_float, _F]
        R.ref_env_eval.argtypes = [C.c_void_p, _F, _F]
        R.ref_env_power.restype = C.c_float
        R.ref_env_power.argtypes = [C.c_void_p]
        R.ref_metrics.argtypes = [_F, _F, C.c_int, C.c_int, C.c_int, C.c_int, _F]
        R.ref_texture_sample.argtypes = [_F, C.c_int, C.c_int, _F, C.c_int, _F]
        R.ref_texture_load.restype = C.c_int
        R.ref_texture_load.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _F]
        R.ref_load_scene.restype = C.c_void_p
        R.ref_load_scene.argtypes = [C.c_char_p]
        R.ref_loaded_sizes.restype = C.c_int
        R.ref_loaded_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        R.ref_loaded_copy.argtypes = [C.c_void_p, _F, _F, _F, _U32, _F, _I32, _F]
        R.ref_loaded_texture_name.restype = C.c_char_p
        R.ref_loaded_texture_name.argtypes = [C.c_void_p, C.c_int]
        R.ref_loaded_free.argtypes = [C.c_void_p]
        R.ref_tonemap.argtypes = [_F, C.c_int, C.c_int, C.c_float, C.c_float, _U8]
        R.ref_write_hdr.restype = C.c_int
        R.ref_write_hdr.argtypes = [C.c_char_p, _F, C.c_int, C.c_int]
        R.ref_full_create.restype = C.c_void_p
        R.ref_full_create.argtypes = [_F, _F, _U32, C.c_uint32, _F, C.c_uint32, _F, C.c_int, C.c_int]
        R.ref_full_destroy.argtypes = [C.c_void_p]
        R.ref_full_lights.restype = C.c_int
        R.ref_full_lights.argtypes = [C.c_void_p, _I32, _I32]
        R.ref_area_light.argtypes = [C.c_void_p, C.c_int, _F, _F, _F]
        R.ref_sample_light.argtypes = [C.c_void_p, _F, C.c_uint32, _I32, _F]
        R.ref_shading_data.argtypes = [C.c_void_p, _F, C.c_uint32, _F]
        R.ref_full_visible.argtypes = [C.c_void_p, _F, _F, C.c_uint32, _U8]
        R.ref_film_splat.argtypes = [C.c_int, C.c_int, _F, _F, C.c_uint32, _F]
        _ref = R
    return _ref


class OracleScene:
    """A scene in the CPU oracle (BVH built on construction)."""

    def __init__(self, data):
        L = lib()
        self._keep = [np.ascontiguousarray(data.positions, np.float32), np.ascontiguousarray(data.normals, np.float32),
                      np.ascontiguousarray(data.material_index, np.uint32),
                      np.ascontiguousarray(data.materials_f32(), np.float32),
                      None if data.env is None else np.ascontiguousarray(data.env, np.float32),
                      data.camera_array()]
        pos, nrm, mi, mats, env, cam = self._keep
        self.width, self.height = data.width, data.height
        self.n_tris = pos.shape[0]
        self.h = L.orc_scene_create(fp(pos), fp(nrm), up(mi), pos.shape[0], fp(mats), mats.shape[0], fp(env),
                                    0 if env is None else env.shape[1], 0 if env is None else env.shape[0], fp(cam))
        if getattr(data, "has_textures", False):
            uv, mt, texels, off, tw, th = data.texture_arrays()
            self._keep += [uv, mt, texels, off, tw, th]
            L.orc_scene_set_textures(self.h, fp(uv), mt.ctypes.data_as(_I32), mt.shape[0], fp(texels),
                                     off.ctypes.data_as(C.POINTER(C.c_uint64)), tw.ctypes.data_as(_I32),
                                     th.ctypes.data_as(_I32), len(data.textures))

    def __del__(self):
        try:
            lib().orc_scene_destroy(self.h)
        except Exception:
            pass

    def bvh(self):
        L = lib()
        n = L.orc_bvh_node_count(self.h)
        nodes = np.empty((n, 8), np.float32)
        idx = np.empty(self.n_tris, np.int32)
        L.orc_bvh_export(self.h, fp(nodes), idx.ctypes.data_as(_I32))
        return nodes, idx

    def max_depth(self) -> int:
        return lib().orc_bvh_max_depth(self.h)

    def camera(self):
        ip, cm, o = np.empty(16, np.float32), np.empty(16, np.float32), np.empty(3, np.float32)
        lib().orc_camera_export(self.h, fp(ip), fp(cm), fp(o))
        return ip, cm, o

    def generate_rays(self, pxy):
        pxy = np.ascontiguousarray(pxy, np.float32).reshape(-1, 2)
        out = np.empty((pxy.shape[0], 6), np.float32)
        lib().orc_generate_rays(self.h, fp(pxy), pxy.shape[0], fp(out))
        return out

    def trace(self, rays, threads=0, counts=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        ids = np.empty(n, np.uint32)
        t, a, b, g = (np.empty(n, np.float32) for _ in range(4))
        cn = np.empty(n, np.uint32) if counts else None
        ct = np.empty(n, np.uint32) if counts else None
        lib().orc_trace_closest(self.h, fp(rays), n, up(ids), fp(t), fp(a), fp(b), fp(g), up(cn), up(ct), threads)
        if counts:
            return ids, t, a, b, g, cn, ct
        return ids, t, a, b, g

    def shadow(self, rays, max_t, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        max_t = np.ascontiguousarray(max_t, np.float32)
        vis = np.empty(rays.shape[0], np.uint8)
        lib().orc_trace_shadow(self.h, fp(rays), fp(max_t), rays.shape[0], vis.ctypes.data_as(_U8), None, None, threads)
        return vis

    def brute(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        ids = np.empty(rays.shape[0], np.uint32)
        t = np.empty(rays.shape[0], np.float32)
        lib().orc_trace_brute(self.h, fp(rays), rays.shape[0], up(ids), fp(t))
        return ids, t

    def render(self, spp, seed=1, spp_offset=0, max_depth=8, rr_depth=3, integrator=0, sampler=0, accum_f64=True,
               threads=0, record=False, env_mode=1, env_visibility=0, crop=None):
        """crop: (x0, y0, width, height) of the pixels traced; the film is full size and
        zero elsewhere. Each pixel keeps its own random stream, so a crop equals the same
        pixels of a full render."""
        p = RenderParams(spp, spp_offset, seed, max_depth, rr_depth, integrator, sampler, 1 if accum_f64 else 0,
                         threads, 1 if record else 0, env_mode, env_visibility, (C.c_int32 * 4)(*(crop or (0, 0, 0, 0))))
        film = np.zeros((self.height, self.width, 3), np.float64)
        st = RenderStats()
        lib().orc_render(self.h, C.byref(p), film.ctypes.data_as(_D), C.byref(st))
        return film, {f: getattr(st, f) for f, _ in RenderStats._fields_}

    def records(self):
        L = lib()
        n = L.orc_records_count(self.h)
        out = np.empty((n, 15), np.float32)
        if n:
            L.orc_records_export(self.h, fp(out))
        return out


def rng(seed, sample, pixel, dim):
    return lib().orc_rng(seed, sample, pixel, dim)


def _ip(a):
    return a.ctypes.data_as(_I32)


class _NeeUnits:
    """Per-unit access to the next-event-estimation constituents; the oracle and the
    reference driver expose the same calls (orc_* / ref_*)."""

    def lights(self):
        n = self._lights(None, None)
        kind, tri = np.empty(n, np.int32), np.empty(n, np.int32)
        self._lights(_ip(kind), _ip(tri))
        return kind, tri

    def area_light(self, light, u2, wi):
        out = np.empty(14, np.float32)
        self._area(int(light), fp(np.ascontiguousarray(u2, np.float32)), fp(np.ascontiguousarray(wi, np.float32)),
                   fp(out))
        return out

    def sample_light(self, u):
        u = np.ascontiguousarray(u, np.float32)
        idx, pmf = np.empty(u.size, np.int32), np.empty(u.size, np.float32)
        self._sample_light(fp(u), u.size, _ip(idx), fp(pmf))
        return idx, pmf

    def shading_data(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = np.empty((rays.shape[0], 24), np.float32)
        self._shading(fp(rays), rays.shape[0], fp(out))
        return out

    def visible(self, p1, p2):
        p1 = np.ascontiguousarray(p1, np.float32).reshape(-1, 3)
        p2 = np.ascontiguousarray(p2, np.float32).reshape(-1, 3)
        vis = np.empty(p1.shape[0], np.uint8)
        self._visible(fp(p1), fp(p2), p1.shape[0], vis.ctypes.data_as(_U8))
        return vis


class OracleNee(_NeeUnits):
    """The oracle's NEE constituents on an OracleScene."""

    def __init__(self, osc):
        L, h = lib(), osc.h
        self.osc = osc
        self._lights = lambda k, t: L.orc_lights(h, k, t)
        self._area = lambda l, u, w, o: L.orc_area_light(h, l, u, w, o)
        self._sample_light = lambda u, n, i, p: L.orc_sample_light(h, u, n, i, p)
        self._shading = lambda r, n, o: L.orc_shading_data(h, r, n, o)
        self._visible = lambda a, b, n, v: L.orc_visible(h, a, b, n, v)


class RefFullScene(_NeeUnits):
    """The reference's own Scene (init + build) over a scene's triangles, materials and
    optional environment map (built only where the reference sources exist)."""

    def __init__(self, data, env=None):
        R = ref()
        self._keep = [np.ascontiguousarray(data.positions, np.float32), np.ascontiguousarray(data.normals, np.float32),
                      np.ascontiguousarray(data.material_index, np.uint32),
                      np.ascontiguousarray(data.materials_f32(), np.float32),
                      None if env is None else np.ascontiguousarray(env, np.float32)]
        pos, nrm, mi, mats, env = self._keep
        h = R.ref_full_create(fp(pos), fp(nrm), up(mi), pos.shape[0], fp(mats), mats.shape[0], fp(env),
                              0 if env is None else env.shape[1], 0 if env is None else env.shape[0])
        self.h = h
        self._lights = lambda k, t: R.ref_full_lights(h, k, t)
        self._area = lambda l, u, w, o: R.ref_area_light(h, l, u, w, o)
        self._sample_light = lambda u, n, i, p: R.ref_sample_light(h, u, n, i, p)
        self._shading = lambda r, n, o: R.ref_shading_data(h, r, n, o)
        self._visible = lambda a, b, n, v: R.ref_full_visible(h, a, b, n, v)

    def __del__(self):
        try:
            ref().ref_full_destroy(self.h)
        except Exception:
            pass


def film_splat(w, h, xy, L):
    """Film::splat with the box filter (the oracle's) of samples in order: (h, w, 3) sums."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    L = np.ascontiguousarray(L, np.float32).reshape(-1, 3)
    film = np.zeros((h, w, 3), np.float32)
    lib().orc_film_splat(w, h, fp(xy), fp(L), xy.shape[0], fp(film))
    return film


def ref_film_splat(w, h, xy, L):
    """Film::splat of the reference's own Film and BoxFilter."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    L = np.ascontiguousarray(L, np.float32).reshape(-1, 3)
    film = np.zeros((h, w, 3), np.float32)
    ref().ref_film_splat(w, h, fp(xy), fp(L), xy.shape[0], fp(film))
    return film
