"""Estimator consistency of the oracle (CPU) on the closed room lit through one window.

Next-event estimation with MIS and plain BSDF sampling are two estimators of the same
image; with the environment shadow ray traced along the light direction their means
agree. The reference aims that shadow ray at sceneCentre + wi * sceneRadius
(Renderer.h:331-332), a point off the light direction for shading points away from the
scene centre: behind a small window it misses the opening and the directly sunlit
objects go black, so that estimator loses about a quarter of the image's energy. Both
renders use fixed seeds of the counter RNG, so the check is deterministic.

The area-light half of next-event estimation is pinned against an estimator that shares
no code with it (tests/independent.py: FP64 numpy, BSDF sampling only); BSDF
reciprocity and the 1/SPP fall of relMSE are checked on the oracle directly.
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


def test_area_light_nee_and_mis_are_unbiased(pg):
    """The oracle's area-light path (computeDirect's area branch, Renderer.h:269-310:
    uniform light pick, Triangle::sample, G = cos cos' / d^2, power-heuristic weight of
    pdf_light * pmf against pdf_bsdf * cos' / d^2; and the MIS-weighted emission of
    BSDF-sampled hits, Renderer.h:536-555) against an FP64 numpy estimator that only
    samples the BSDF and adds emission where a path hits a light (tests/independent.py),
    on a closed room with two differently sized and coloured area lights and an
    occluder (max depth 64, so the depth cap's missing BSDF half at the last vertex
    is below 1e-9). Every 8 x 8 pixel block sum agrees within 4.5 standard errors and
    the image total within 1%. Planted in a CPU copy of the oracle, a BSDF area pdf
    without the 1/d^2 moves a block by 8.8 standard errors, a light pdf without the
    pmf by 120, an emission weight with the pdfs exchanged by 73 and a geometry term
    without cos' by 179."""
    from oracle.pg_oracle import OracleScene
    import independent as ind

    def render(data, seed):
        film, _ = OracleScene(data).render(64, seed=seed, max_depth=64, accum_f64=True)
        return ind.block_sums(film / 64)

    z, total = ind.against_fixture(pg, render)
    assert np.abs(z).max() < 4.5, np.round(z, 2)
    assert abs(total - 1.0) < 0.01, total


def test_bsdf_reciprocity_of_oracle():
    """evaluate(wo -> wi) = evaluate(wi -> wo) for the reciprocal BSDFs (Diffuse,
    rough GGX Conductor, Oren-Nayar), on both sides of the surface."""
    from oracle import pg_oracle as po
    rng = np.random.default_rng(5)
    n = np.array([0.2, 0.9, -0.3])
    n /= np.linalg.norm(n)
    checked = 0
    for kind, rough in ((0, 0.0), (2, 0.3), (2, 0.8), (5, 0.4)):
        mat = np.zeros(16, np.float32)
        mat[0], mat[1:4], mat[7:10], mat[10:13], mat[15] = kind, (0.7, 0.5, 0.3), (0.2, 0.9, 1.1), (3.9, 2.4, 2.1), rough
        for _ in range(200):
            a, b = rng.normal(size=3), rng.normal(size=3)
            a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
            a, b = a * np.sign(np.dot(a, n)), b * np.sign(np.dot(b, n))  # both above the surface
            if np.dot(a, n) < 0.05 or np.dot(b, n) < 0.05:
                continue
            f = []
            for wo, wi in ((a, b), (b, a)):
                sd = np.zeros(12, np.float32)
                sd[3:6], sd[6:9], sd[9:12] = wo, n, n
                out, pdf = np.empty(3, np.float32), np.empty(1, np.float32)
                po.lib().orc_bsdf_eval(po.fp(np.ascontiguousarray(mat)), po.fp(sd),
                                       po.fp(np.ascontiguousarray(wi, np.float32)), po.fp(out), po.fp(pdf))
                f.append(out.astype(np.float64))
            assert np.allclose(f[0], f[1], rtol=2e-4, atol=1e-7), (kind, rough, f)
            checked += 1
    assert checked > 300


def test_relmse_falls_as_one_over_spp(pg):
    """The Cornell box's relMSE against the FP64 4096-SPP oracle image scales as 1/SPP
    (an unbiased estimator's variance): relMSE(4) / relMSE(64) is 16 within the noise
    of the 4-SPP images (checked over four seeds each)."""
    from oracle.pg_oracle import OracleScene
    from conftest import GOLDEN
    ref = np.load(GOLDEN / "oracle_config1_4096spp.npz")["image"].astype(np.float64)
    osc = OracleScene(pg.generate_scene(1, 0))
    r = {}
    for spp in (4, 64):
        vals = []
        for s in range(4):
            film, _ = osc.render(spp, seed=500 + s, accum_f64=True)
            vals.append(pg.relmse(film / spp, ref))
        r[spp] = float(np.mean(vals))
    ratio = r[4] / r[64]
    assert 11.0 < ratio < 21.0, (r, ratio)
