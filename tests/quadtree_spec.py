"""The directional quadtree of one guide cell built from its specification, for the
node-for-node check of the device build (guiding.hip k_qt_build): directions map to
(u, v) in [0, 1)^2 by the equal-area cylindrical map (u = phi / 2 pi with
phi = atan2(y, x) in [0, 2 pi), v = (z + 1) / 2); a node splits its square into four
quadrants (index qx + 2 qy); a quadrant becomes a child node when its flux exceeds
1% of the cell's total flux and the child would sit above depth 10 (root depth 0,
so the finest quadrants are 2^-10 wide). Flux is the record's luminance
(0.2126 R + 0.7152 G + 0.0722 B) in 64-bit fixed point (2^24 units, rounded), so
sums do not depend on order.

TEST INFRASTRUCTURE ONLY.
"""
from __future__ import annotations

import numpy as np

MAX_DEPTH = 10
SCALE = float(1 << 24)


def fixed_flux(lum_f32: np.ndarray) -> np.ndarray:
    return np.floor(lum_f32.astype(np.float64) * SCALE + 0.5).astype(np.uint64)


def build(u: np.ndarray, v: np.ndarray, flux: np.ndarray):
    """Nested tree: (energies[4] as uint64, children[4] as subtree or None)."""
    total = int(flux.sum(dtype=np.uint64))

    def node(sel, ox, oy, size, depth):
        half = size * 0.5
        q = (u[sel] >= ox + half).astype(int) + 2 * (v[sel] >= oy + half).astype(int)
        e = [int(flux[sel][q == k].sum(dtype=np.uint64)) for k in range(4)]
        kids = []
        for k in range(4):
            if depth + 1 < MAX_DEPTH and total > 0 and e[k] * 100 > total:
                kids.append(node(sel[q == k], ox + (k & 1) * half, oy + (k >> 1) * half, half, depth + 1))
            else:
                kids.append(None)
        return e, kids

    return node(np.arange(u.size), 0.0, 0.0, 1.0, 0)


def compare(spec, nodes: np.ndarray, at: int = 0, path: str = "root"):
    """Walk the spec tree and an exported pool (8 words per node: 4 flux floats, 4
    child indices) together; returns the number of nodes compared or raises."""
    e, kids = spec
    E = nodes[at, :4]
    ch = nodes[at, 4:8].view(np.int32)
    want = (np.array(e, np.float64) / SCALE).astype(np.float32)
    if not np.array_equal(E, want):
        raise AssertionError(f"{path}: energies {E} != {want}")
    count = 1
    for k in range(4):
        if (kids[k] is None) != (ch[k] < 0):
            raise AssertionError(f"{path}/{k}: refined {ch[k] >= 0}, spec {kids[k] is not None}")
        if kids[k] is not None:
            count += compare(kids[k], nodes, int(ch[k]), f"{path}/{k}")
    return count
