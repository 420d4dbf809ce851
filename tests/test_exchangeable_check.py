"""The sign test of tests/exchangeable.py on the oracle (CPU): two renders of the same
estimator with different seeds pass it, and a 5% bias planted in the conductor or the
glass BSDF of one render (the same random stream otherwise) fails it, with that class's
sum leaning to the planted sign."""
from dataclasses import replace

import numpy as np
import pytest

import exchangeable as ex


def _setup(pg):
    from oracle.pg_oracle import OracleScene
    base = pg.generate_scene(100, 0)
    data = replace(base, camera=dict(base.camera, width=96, height=96))
    osc = OracleScene(data)
    ys, xs = np.mgrid[0:96, 0:96]
    pxy = np.stack([xs.reshape(-1) + 0.5, ys.reshape(-1) + 0.5], 1).astype(np.float32)
    labels = ex.classes_of_hits(data, osc.trace(osc.generate_rays(pxy))[0]).reshape(96, 96)
    return data, osc, labels


def _render(osc, seed):
    film, _ = osc.render(4, seed=seed, max_depth=12, accum_f64=False, env_mode=1, threads=8)
    return film


def test_independent_seeds_pass(pg):
    data, osc, labels = _setup(pg)
    a = _render(osc, 3)
    for seed in (4, 5, 6):
        ex.assert_exchangeable(_render(osc, seed), a, labels)
    assert ex.z_by_class(a, a, labels)["all"] == 0.0


def test_planted_bsdf_bias_fails_in_its_class(pg):
    from oracle.pg_oracle import OracleScene
    data, osc, labels = _setup(pg)
    a = _render(osc, 3)
    for bsdf, name in ((2, "bsdf2"), (3, "bsdf3")):
        m = data.materials.copy()
        m["albedo"][m["type"] == bsdf] *= 0.95
        b = _render(OracleScene(replace(data, materials=m)), 3)
        with pytest.raises(AssertionError):
            ex.assert_exchangeable(b, a, labels)
        z = ex.z_by_class(b, a, labels)
        assert z[name] < -3.0 and z["all"] < -4.5, (name, z)
