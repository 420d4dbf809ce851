"""Estimator consistency of the oracle (CPU) on the closed room lit through one window.

Next-event estimation with MIS and plain BSDF sampling are two estimators of the same
image; with the environment shadow ray traced along the light direction their means
agree. The reference aims that shadow ray at sceneCentre + wi * sceneRadius
(Renderer.h:331-332), a point off the light direction for shading points away from the
scene centre: behind a small window it misses the opening and the directly sunlit
objects go black, so that estimator loses about a quarter of the image's energy. Both
renders use fixed seeds of the counter RNG, so the check is deterministic.
"""
from dataclasses import replace

import numpy as np


def test_env_visibility_along_light_direction_is_consistent(pg):
    from oracle.pg_oracle import OracleScene
    base = pg.generate_scene(101, 0)
    data = replace(base, camera=dict(base.camera, width=90, height=160))
    osc = OracleScene(data)
    spp = 256
    mean = {}
    for name, kw in (("reference", dict(integrator=0, env_visibility=0)),
                     ("directional", dict(integrator=0, env_visibility=1)),
                     ("bsdf_only", dict(integrator=1))):
        film, _ = osc.render(spp, seed=3, max_depth=12, threads=8, **kw)
        mean[name] = float(film.mean() / spp)
    assert abs(mean["directional"] / mean["bsdf_only"] - 1.0) < 0.05, mean
    assert mean["reference"] / mean["bsdf_only"] < 0.85, mean
    assert np.isfinite(list(mean.values())).all()
