import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def pg():
    import pathguidingdissertation_amd as pg
    pg.load()
    return pg


@pytest.fixture(scope="session")
def device(pg):
    dev = pg.Device(0)
    yield dev
    dev.close()


@pytest.fixture(scope="session")
def cornell(pg):
    return pg.generate_scene(1, 0)


@pytest.fixture(scope="session")
def small_interior(pg):
    return pg.generate_scene(100, 0)


def clustered_scene(pg):
    """Cornell box plus two clusters whose triangles share one centroid each (40
    triangles rotated about a common centre, 100 exact copies of one triangle): no
    split separates equal centroids, so the BVH holds leaves of 40 and 100 triangles,
    beyond the 30 that a traversal stack entry stores inline."""
    from dataclasses import replace
    base = pg.generate_scene(1, 0)
    tris = []
    c = np.array([0.2, 0.8, -0.1])
    for k in range(40):
        a = np.pi * k / 40
        u = np.array([np.cos(a), np.sin(a), 0.3 * np.sin(3 * a)])
        w = np.cross(u, [0.3, 0.4, 0.87])
        w /= np.linalg.norm(w)
        v = np.stack([c + 0.3 * u, c - 0.15 * u + 0.25 * w, c - 0.15 * u - 0.25 * w])
        v -= v.mean(0) - c
        tris.append(v)
    one = np.array([[-0.5, 1.2, 0.3], [-0.2, 1.5, 0.1], [-0.4, 1.3, 0.6]])
    tris += [one] * 100
    extra = np.asarray(tris, np.float32)
    n_extra = extra.shape[0]
    nrm = np.cross(extra[:, 1] - extra[:, 0], extra[:, 2] - extra[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm3 = np.repeat(nrm[:, None, :], 3, 1).astype(np.float32)
    return replace(base, positions=np.concatenate([base.positions, extra]),
                   normals=np.concatenate([base.normals, nrm3]),
                   material_index=np.concatenate([base.material_index, np.zeros(n_extra, np.uint32)]))


def seeded_rays(osc, data, n, seed):
    """Seeded parity ray set: half camera primaries with jitter, half uniform-sphere
    rays from on-surface origins offset by 1e-4 along the normal (Renderer.h:496-497),
    plus axis-aligned directions so that 1/d = +-inf is exercised."""
    rng = np.random.default_rng(seed)
    n_cam = n // 2
    pxy = np.stack([rng.uniform(0, data.width, n_cam), rng.uniform(0, data.height, n_cam)], 1).astype(np.float32)
    cam = osc.generate_rays(pxy)
    m = n - n_cam
    tri = rng.integers(0, data.n_tris, m)
    r1, r2 = rng.random(m), rng.random(m)
    s = np.sqrt(r1)
    a, b = 1 - s, r2 * s
    g = 1 - a - b
    P = data.positions[tri].astype(np.float64)
    org = a[:, None] * P[:, 0] + b[:, None] * P[:, 1] + g[:, None] * P[:, 2]
    nrm = np.cross(P[:, 2] - P[:, 1], P[:, 0] - P[:, 2])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    z = rng.uniform(-1, 1, m)
    phi = rng.uniform(0, 2 * np.pi, m)
    d = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z], 1)
    k = m // 16
    axis = rng.integers(0, 3, k)
    sign = rng.choice([-1.0, 1.0], k)
    d[:k] = 0.0
    d[np.arange(k), axis] = sign
    side = np.sign(np.sum(d * nrm, 1))
    side[side == 0] = 1
    org = org + nrm * (1e-4 * side)[:, None]
    sph = np.concatenate([org, d], 1).astype(np.float32)
    return np.concatenate([cam, sph], 0)
