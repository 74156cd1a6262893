"""The hand-made deep trees (tests/deep_trees.py) are valid trees in the
reference's upload format: the oracle's stackless preorder walk of them
(pathtrace.metal:119-163) finds exactly the brute-force closest hits."""
import numpy as np
import pytest

import deep_trees
import oracle_bindings as ob


@pytest.mark.parametrize("make,depth", [(deep_trees.comb, 120), (deep_trees.caterpillar, 20000)],
                         ids=["comb120", "caterpillar20000"])
def test_deep_tree_fixture_matches_brute_force(make, depth):
    e = make(depth)
    o, d = deep_trees.rays(depth, 2048, 7)
    tri, t, _, _, _ = ob.trace(e, o, d)
    btri, bt = ob.trace_brute(e, o, d)
    assert (tri == btri).all()
    assert (t[tri >= 0] == bt[tri >= 0]).all()
    # both kinds of rays occur: hits and rays that pass every plane
    assert 0.2 < np.mean(tri >= 0) < 0.9
