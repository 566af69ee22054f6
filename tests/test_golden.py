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

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_*.npz"))
                if not p.endswith("reference_clip.npz"))


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


def test_builder_clip_fast_path_matches_reference_golden():
    """The builder's register-only clip (only the two planes of the split axis cut) against
    the compiled reference's clipTriangleToAABB on the committed cases: bit-identical boxes
    wherever it applies, the empty box where the reference's polygon vanishes, and it
    declines the cases where a plane of another axis cuts."""
    import rtx
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_clip.npz"))
    v, reg, axis, box, valid = g["verts"], g["region"], g["axis"], g["box"], g["valid"]
    taken = 0
    for i in range(len(v)):
        ok, b = rtx.debug_clip_slab(v[i], reg[i], int(axis[i]))
        p = v[i].reshape(3, 3)
        others = [a for a in range(3) if a != axis[i]]
        cut_elsewhere = any(p[:, a].min() < reg[i, a] or p[:, a].max() > reg[i, 3 + a] for a in others)
        if cut_elsewhere and not any(axis[i] < a for a in others if p[:, a].min() < reg[i, a] or p[:, a].max() > reg[i, 3 + a]):
            assert not ok, i   # a plane before the split axis cuts the triangle itself
        if not ok:
            continue
        taken += 1
        if valid[i]:
            assert np.array_equal(b.view(np.uint32), box[i].view(np.uint32)), (i, b, box[i])
        else:
            assert not (b[0] <= b[3] and b[1] <= b[4] and b[2] <= b[5]), (i, b)
    assert taken > len(v) // 2
