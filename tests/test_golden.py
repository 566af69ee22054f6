"""The oracle against committed outputs of the compiled reference (tests/golden/, written
by tests/golden/make_golden.py from oracle/_ref): Morton codes, scene box, sort order,
builder-order and DFS-order nodes, reordered triangles and ids, index-array sizes, and
closest-hit t/u/v/triIndex/normal on random and edge-case rays, all bit for bit. These
hold without the reference checkout."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from scenes_util import assert_same_nodes

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_*.npz")))


def test_golden_files_present():
    assert len(GOLDEN) >= 8


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[10:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(path):
    g = np.load(path)
    v, t = g["verts"], g["tris"]
    codes, box = O.morton(v, t)
    assert np.array_equal(codes, g["codes"])
    assert np.array_equal(box.view(np.uint32), g["scene_box"].view(np.uint32))
    o = O.build(v, t)
    assert np.array_equal(o.sorted, g["sorted"])
    name = os.path.basename(path)
    assert_same_nodes(o.nodes, np.ascontiguousarray(g["nodes"]).view(O.NODE48).reshape(-1), name + " DFS order")
    assert_same_nodes(o.bfs_nodes, np.ascontiguousarray(g["bfs_nodes"]).view(O.NODE48).reshape(-1),
                      name + " builder order")
    assert np.array_equal(o.tris, g["rtris"])
    assert np.array_equal(o.prims, g["prims"])
    assert [o.info["tindex_size"], o.info["sorted_size"]] == list(g["sizes"])
    for key in ("rays", "edge"):
        h = O.trace(o, g[key], normals=True)
        ref = g[key + "_hit"]
        got = np.stack([h["t"], h["u"], h["v"], h["tri"].view(np.float32)], 1)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), key
        assert np.array_equal(h["normal"].view(np.uint32), g[key + "_normal"].view(np.uint32)), key
