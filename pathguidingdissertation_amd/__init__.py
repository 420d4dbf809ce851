"""MI355X-native path-tracing hot path of IsaGeriler/PathGuidingDissertation.

The render loop, BVH traversal, shading and film run as hand-written gfx950 kernels
in ``libpgtrace.so`` behind the C ABI of ``include/pg_tracer.h``; this package is the
thin host-side mirror of the reference's Scene / RayTracer interfaces.
"""
from ._lib import LIB_PATH, PgError, load  # noqa: F401
from .scene import (  # noqa: F401
    BSDF_CONDUCTOR, BSDF_DIELECTRIC, BSDF_DIFFUSE, BSDF_GLASS, BSDF_LAYERED, BSDF_MIRROR, BSDF_ORENNAYAR,
    BSDF_PLASTIC, Batch, Device, RayTracer, Scene, SceneData, generate_scene, load_scene, metrics, read_image,
    relmse, write_hdr, write_png,
)
