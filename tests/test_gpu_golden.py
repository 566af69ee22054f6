"""The GPU builder and traversal against committed outputs of the compiled reference
(tests/golden/): nodes, triangles, original ids and closest hits, bit for bit."""
import glob
import os

import numpy as np
import pytest

import rtx
from oracle import oracle as O
from scenes_util import assert_same_nodes

pytestmark = pytest.mark.gpu

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_*.npz"))
                if not p.endswith("reference_clip.npz"))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[10:-4] for p in GOLDEN])
def test_gpu_matches_reference_golden(path):
    g = np.load(path)
    print(f"[golden] {os.path.basename(path)}", flush=True)
    b = rtx.build(g["verts"], g["tris"])
    nodes, tris, prims = b.export()
    name = os.path.basename(path)
    assert_same_nodes(nodes, np.ascontiguousarray(g["nodes"]).view(O.NODE48).reshape(-1), name + " DFS order")
    assert_same_nodes(b.export_bfs(), np.ascontiguousarray(g["bfs_nodes"]).view(O.NODE48).reshape(-1),
                      name + " builder order")
    assert np.array_equal(tris, g["rtris"]) and np.array_equal(prims, g["prims"])
    codes, srt = b.morton()
    assert np.array_equal(codes, g["codes"]) and np.array_equal(srt, g["sorted"])
    for key in ("rays", "edge"):
        h = b.trace(g[key], normals=True)
        ref = g[key + "_hit"]
        rtri = ref[:, 3].view(np.uint32)
        assert np.array_equal(h["tri"], rtri), key
        hit = rtri != 0xFFFFFFFF
        got = np.stack([h["t"], h["u"], h["v"]], 1)
        assert np.array_equal(got[hit].view(np.uint32), ref[hit, :3].view(np.uint32)), key
        np.testing.assert_allclose(h["normal"][hit], g[key + "_normal"][hit], rtol=0, atol=1e-6)
