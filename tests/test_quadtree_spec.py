"""The quadtree specification builder itself (tests/quadtree_spec.py) on cases whose
trees are known by hand."""
import numpy as np

import quadtree_spec as qs


def test_point_flux_refines_to_depth_ten():
    # all flux in one direction: one chain of nodes down to depth 9, one child each
    u = np.full(10, 0.3)
    v = np.full(10, 0.7)
    tree = qs.build(u, v, np.full(10, 1 << 20, np.uint64))
    depth, node = 0, tree
    while True:
        e, kids = node
        assert sum(e) == 10 << 20 and sum(x > 0 for x in e) == 1
        refined = [k for k in kids if k is not None]
        if not refined:
            break
        assert len(refined) == 1
        node, depth = refined[0], depth + 1
    assert depth == qs.MAX_DEPTH - 1


def test_one_percent_threshold_is_strict():
    # quadrant flux exactly 1% of the total is not refined; just above it is
    u = np.array([0.1, 0.9, 0.9])
    v = np.array([0.1, 0.9, 0.9])
    flux = np.array([1, 50, 49], np.uint64)  # total 100: quadrant 0 holds exactly 1%
    e, kids = qs.build(u, v, flux)
    assert e == [1, 0, 0, 99] and kids[0] is None and kids[3] is not None
    flux = np.array([2, 50, 49], np.uint64)  # 2 of 101 > 1%
    e, kids = qs.build(u, v, flux)
    assert kids[0] is not None
