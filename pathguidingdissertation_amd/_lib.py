"""ctypes binding of the C ABI declared in include/pg_tracer.h.

The product path is the gfx950 library ``libpgtrace.so``; there is no CPU fallback.
Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libpgtrace.so"

PG_OK = 0
PG_ERR_INVALID = -1
PG_ERR_NOMEM = -2
PG_ERR_DEVICE = -3
PG_ERR_OVERFLOW = -4
PG_ERR_STATE = -5

PG_INTEGRATOR_PATH = 0
PG_INTEGRATOR_RECORD = 1
PG_INTEGRATOR_GUIDED_QT = 2
PG_INTEGRATOR_GUIDED_MLP = 3
PG_INTEGRATOR_GUIDED_PSS = 4

PG_LOAD_SKIP_MISSING = 1


class PgMaterial(C.Structure):
    _fields_ = [("type", C.c_float), ("albedo", C.c_float * 3), ("emission", C.c_float * 3),
                ("eta", C.c_float * 3), ("k", C.c_float * 3), ("int_ior", C.c_float),
                ("ext_ior", C.c_float), ("roughness", C.c_float)]


class PgCameraDesc(C.Structure):
    _fields_ = [("from_", C.c_float * 3), ("to", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("width", C.c_float), ("height", C.c_float)]


class PgRenderParams(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("integrator", C.c_int32), ("record", C.c_uint32), ("guide_prob", C.c_float),
                ("knn_k", C.c_uint32), ("record_nee", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("sort_rays", C.c_uint32),
                ("dyn_fetch", C.c_uint32), ("env_visibility", C.c_uint32), ("samples_per_pass", C.c_uint32),
                ("stack_lds", C.c_uint32), ("guide_depth", C.c_uint32)]


class PgStats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("records", C.c_uint64), ("kernel_ms_extend", C.c_double), ("kernel_ms_shade", C.c_double),
                ("kernel_ms_shadow", C.c_double), ("kernel_ms_total", C.c_double),
                ("extend_launches", C.c_uint64), ("shadow_launches", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("tris_tested", C.c_uint64),
                ("shadow_nodes_visited", C.c_uint64), ("shadow_tris_tested", C.c_uint64),
                ("kernel_ms_guide", C.c_double)]


class PgMlpConfig(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("log2_table", C.c_uint32), ("res_min", C.c_float), ("res_max", C.c_float),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("batch", C.c_uint32), ("weight_clip", C.c_float), ("act_budget", C.c_uint64),
                ("nan_step", C.c_uint32), ("diverge_rise", C.c_float), ("diverge_patience", C.c_uint32),
                ("rise_test_step", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class PgMlpInfo(C.Structure):
    _fields_ = [("ready", C.c_uint32), ("diverged", C.c_uint32), ("step", C.c_uint32), ("n_params", C.c_uint32),
                ("n_grid_params", C.c_uint64), ("last_loss", C.c_float), ("w_scale", C.c_float),
                ("chunks", C.c_uint32), ("train_ms", C.c_float), ("fallbacks", C.c_uint32),
                ("diverge_reason", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class PgMlpSchedule(C.Structure):
    _fields_ = [("train_every", C.c_uint32), ("steps", C.c_uint32), ("window", C.c_uint32),
                ("warmup_passes", C.c_uint32), ("warmup_steps", C.c_uint32), ("record_samples", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class PgImageMetrics(C.Structure):
    _fields_ = [("mse", C.c_float), ("rmse", C.c_float), ("rrmse", C.c_float), ("relmse", C.c_double)]


class PgBatchConfig(C.Structure):
    _fields_ = [("queue_paths", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32), ("shade_blocks_per_cu", C.c_uint32),
                ("stack", C.c_uint32), ("use_graph", C.c_uint32), ("dyn_fetch", C.c_uint32),
                ("samples_per_pass", C.c_uint32), ("stack_lds", C.c_uint32)]


class PgBatchStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("gpu_ms", C.c_double), ("views", C.c_uint32), ("scenes", C.c_uint32), ("chunks", C.c_uint32),
                ("graph_nodes", C.c_uint32)]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "pg_create": (C.c_int32, [C.POINTER(_P), C.c_int]),
    "pg_destroy": (C.c_int32, [_P]),
    "pg_last_error": (C.c_char_p, [_P]),
    "pg_device_info": (C.c_int32, [_P, C.c_char_p, C.c_size_t, _I32, C.POINTER(C.c_double)]),
    "pg_scene_upload": (C.c_int32, [_P, _F, _F, _U32, C.c_uint32, C.POINTER(PgMaterial), C.c_uint32, _F,
                                    C.c_int32, C.c_int32, C.POINTER(PgCameraDesc)]),
    "pg_bvh_info": (C.c_int32, [_P, _U32, _U32, C.POINTER(C.c_double)]),
    "pg_bvh_export": (C.c_int32, [_P, _F, _I32]),
    "pg_trace_closest": (C.c_int32, [_P, _F, C.c_uint32, _P]),
    "pg_trace_closest_noalloc": (C.c_int32, [_P, _F, C.c_uint32, _P]),
    "pg_bvh_build_gpu": (C.c_int32, [_P, _F, C.c_uint32, _U32, _F, _I32, C.POINTER(C.c_double)]),
    "pg_set_bvh_builder": (C.c_int32, [_P, C.c_int32]),
    "pg_trace_shadow": (C.c_int32, [_P, _F, _F, C.c_uint32, _U8]),
    "pg_generate_rays": (C.c_int32, [_P, _F, C.c_uint32, _F]),
    "pg_rng_values": (C.c_int32, [_P, C.c_uint32, C.c_uint32, _U32, _U32, C.c_uint32, _F]),
    "pg_bsdf_query": (C.c_int32, [_P, C.POINTER(PgMaterial), C.c_uint32, _F, _F, _F, _F]),
    "pg_render_params_default": (PgRenderParams, []),
    "pg_film_clear": (C.c_int32, [_P]),
    "pg_render_spp": (C.c_int32, [_P, C.POINTER(PgRenderParams), C.c_uint32]),
    "pg_film_read": (C.c_int32, [_P, _F, _U32]),
    "pg_render_tune": (C.c_int32, [_P, C.POINTER(PgRenderParams), C.POINTER(PgRenderParams),
                                   C.POINTER(C.c_double)]),
    "pg_stats_read": (C.c_int32, [_P, C.POINTER(PgStats)]),
    "pg_stats_reset": (C.c_int32, [_P]),
    "pg_set_counting": (C.c_int32, [_P, C.c_int32]),
    "pg_set_timing": (C.c_int32, [_P, C.c_int32]),
    "pg_copy_benchmark": (C.c_int32, [_P, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "pg_records_count": (C.c_int32, [_P, C.POINTER(C.c_uint64)]),
    "pg_records_read": (C.c_int32, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "pg_records_clear": (C.c_int32, [_P]),
    "pg_records_upload": (C.c_int32, [_P, _P, C.c_uint64]),
    "pg_cache_build": (C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    "pg_cache_info": (C.c_int32, [_P, C.POINTER(C.c_uint64), _U32, _U32, _U32, C.POINTER(C.c_double)]),
    "pg_cache_export": (C.c_int32, [_P, _F, _F, _F]),
    "pg_knn": (C.c_int32, [_P, _F, C.c_uint32, C.c_uint32, _U32, _F]),
    "pg_guide_qt_eval": (C.c_int32, [_P, _F, _F, _F, C.c_uint32, _F, _F, _F]),
    "pg_guide_pss_eval": (C.c_int32, [_P, C.c_uint32, _F, _F, _F, _F, _F, C.c_uint32, _F, _F, _F]),
    "pg_mlp_config_default": (PgMlpConfig, []),
    "pg_mlp_init": (C.c_int32, [_P, C.POINTER(PgMlpConfig)]),
    "pg_mlp_train": (C.c_int32, [_P, C.c_uint32, _F]),
    "pg_mlp_render_online": (C.c_int32, [_P, C.POINTER(PgRenderParams), C.c_uint32, C.POINTER(PgMlpSchedule)]),
    "pg_mlp_info_read": (C.c_int32, [_P, C.POINTER(PgMlpInfo)]),
    "pg_mlp_get_params": (C.c_int32, [_P, _F, _F]),
    "pg_mlp_set_params": (C.c_int32, [_P, _F, _F]),
    "pg_mlp_levels": (C.c_int32, [_P, _U32, _U32, _F]),
    "pg_mlp_eval": (C.c_int32, [_P, C.c_uint32, _F, _F, _F, _F, _F, _F, _F, _F]),
    "pg_mlp_gradient": (C.c_int32, [_P, C.c_uint32, _F, _F, _F, _F]),
    "pg_batch_create": (C.c_int32, [_P, C.POINTER(_P)]),
    "pg_batch_destroy": (C.c_int32, [_P]),
    "pg_batch_add_scene": (C.c_int32, [_P, _F, _F, _U32, C.c_uint32, C.POINTER(PgMaterial), C.c_uint32, _F,
                                       C.c_int32, C.c_int32, _U32]),
    "pg_batch_scene_textures": (C.c_int32, [_P, C.c_uint32, _F, _I32, C.c_uint32, _F, C.POINTER(C.c_uint64), _I32,
                                            _I32, C.c_uint32]),
    "pg_batch_add_view": (C.c_int32, [_P, C.c_uint32, C.POINTER(PgCameraDesc), _U32]),
    "pg_batch_configure": (C.c_int32, [_P, C.POINTER(PgBatchConfig)]),
    "pg_batch_autotune": (C.c_int32, [_P, C.POINTER(PgRenderParams), C.POINTER(PgBatchConfig),
                                      C.POINTER(C.c_double)]),
    "pg_batch_clear": (C.c_int32, [_P]),
    "pg_batch_render": (C.c_int32, [_P, C.POINTER(PgRenderParams), C.c_uint32]),
    "pg_batch_set_guide": (C.c_int32, [_P, C.c_uint32, C.POINTER(PgMlpConfig), _F, _F]),
    "pg_batch_film_read": (C.c_int32, [_P, C.c_uint32, _F, _U32]),
    "pg_batch_stats_read": (C.c_int32, [_P, C.POINTER(PgBatchStats)]),
    "pg_metrics": (C.c_int32, [_F, _F, C.c_uint64, C.c_uint64, C.c_int32, C.c_int32, C.POINTER(PgImageMetrics)]),
    "pg_scene_generate": (C.c_int32, [C.c_int32, C.c_uint32, C.POINTER(_P)]),
    "pg_scene_sizes": (C.c_int32, [_P, _U32, _U32, _I32, _I32]),
    "pg_scene_copy": (C.c_int32, [_P, _F, _F, _U32, C.POINTER(PgMaterial), _F, C.POINTER(PgCameraDesc)]),
    "pg_scene_free": (C.c_int32, [_P]),
    "pg_bvh_build_host": (C.c_int32, [_F, C.c_uint32, _U32, _F, _I32]),
    "pg_scene_textures": (C.c_int32, [_P, _F, _I32, C.c_uint32, _F, C.POINTER(C.c_uint64), _I32, _I32,
                                      C.c_uint32]),
    "pg_scene_load": (C.c_int32, [C.c_char_p, C.c_uint32, C.POINTER(_P)]),
    "pg_scene_load_error": (C.c_char_p, []),
    "pg_scene_texture_sizes": (C.c_int32, [_P, _U32, C.POINTER(C.c_uint64), _U32]),
    "pg_scene_copy_textures": (C.c_int32, [_P, _F, _I32, _F, C.POINTER(C.c_uint64), _I32, _I32]),
    "pg_scene_warning": (C.c_char_p, [_P, C.c_uint32]),
    "pg_image_read": (C.c_int32, [C.c_char_p, _I32, _I32, _I32, _F]),
    "pg_image_write_png": (C.c_int32, [C.c_char_p, _U8, C.c_int32, C.c_int32]),
    "pg_image_write_hdr": (C.c_int32, [C.c_char_p, _F, C.c_int32, C.c_int32]),
    "pg_film_tonemap": (C.c_int32, [_P, C.c_float, C.c_float, _U8]),
}

_lib = None


def load() -> C.CDLL:
    """Load libpgtrace.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (the HIP path has no fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pg error {code}: {msg}")
        self.code = code


def check(code: int, ctx=None) -> None:
    if code != PG_OK:
        lib = load()
        msg = lib.pg_last_error(ctx)
        raise PgError(code, msg.decode() if msg else "")
