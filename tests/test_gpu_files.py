"""The .bvh file format and the bake -> load -> trace flow of the reference's
tests/test_traverse.cpp, through the GPU library."""
import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S
from scenes_util import assert_same_nodes

pytestmark = pytest.mark.gpu


def _write_reference_layout(path, nodes, tris, verts):
    hdr = np.array([1, len(nodes), len(tris), len(verts)], np.uint32)
    path.write_bytes(hdr.tobytes() + nodes.tobytes() + np.ascontiguousarray(tris, np.uint32).tobytes()
                     + np.ascontiguousarray(verts, np.float32).tobytes())


def test_save_load_round_trip(tmp_path):
    v, t = S.random_triangles(3000, seed=21)
    g = rtx.build(v, t)
    g.save(tmp_path / "a.bvh")
    h = rtx.load_bvh(tmp_path / "a.bvh")
    n1, t1, _ = g.export()
    n2, t2, _ = h.export()
    assert_same_nodes(n1, n2, "round trip")
    assert np.array_equal(t1, t2)
    rays = S.incoherent_rays(5000, seed=2)
    a, b = g.trace(rays), h.trace(rays)
    assert np.array_equal(a["tri"], b["tri"]) and np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


def test_loads_a_tree_in_the_reference_layout(tmp_path):
    """A tree in the reference's on-disk layout (writeBakedData, bvh_IO.cpp:257-264) loads
    and traces like the reference's own trace()."""
    v, t = S.random_triangles(4000, seed=8)
    tree = O.ref_build(v, t) if O.ref_available() else O.build(v, t)
    _write_reference_layout(tmp_path / "r.bvh", tree.nodes, tree.tris, v)
    h = rtx.load_bvh(tmp_path / "r.bvh")
    rays = S.incoherent_rays(8000, seed=3)
    got = h.trace(rays, normals=True)
    exp = O.ref_trace(tree, rays, normals=True) if O.ref_available() else O.trace(tree, rays, normals=True)
    assert np.array_equal(got["tri"], exp["tri"])
    assert np.array_equal(got["t"].view(np.uint32), exp["t"].view(np.uint32))
    np.testing.assert_allclose(got["normal"], exp["normal"], atol=1e-6)


def test_bake_load_trace_flow(tmp_path):
    """bakeBVH(scene) -> loadBVH -> trace, on the scene of tests/test_traverse.cpp:73-76."""
    scn = tmp_path / "test_scene.scn"
    scn.write_text("box 0 0 5 1 1 1 q 0 0 0 1\nbox 0 0 12 1 1 1 q 0 0 0 1\nsphere 0 5 0 2\n"
                   "box 0 -3 0 20 0.1 20 q 0 0 0 1\n")
    rtx.bake_bvh(scn)
    h = rtx.load_bvh(tmp_path / "test_scene.bvh")
    v, t = rtx.load_scene(scn)
    o = O.build(v, t)
    nodes, tris, _ = h.export()
    assert_same_nodes(nodes, o.nodes, "baked file")
    assert np.array_equal(tris, o.tris)
    ray = np.array([[0, -5, 0, 0, 0, 1, 0, np.inf]], np.float32)   # ray 11: ground through spatial splits
    hit = h.trace(ray, normals=True)
    assert hit["tri"][0] != 0xFFFFFFFF and 1.9 <= hit["t"][0] <= 2.2 and hit["normal"][0, 1] > 0


def test_ply_mesh_load_build_trace_render(tmp_path):
    """Mesh ingest on the GPU path: the PLY-derived mesh of reference_bunny_head.npz (a
    `mesh` line of a .scn file, bvh_IO.cpp:195-231, read by rtx_scene_load in place of the
    assimp import of loadMesh, bvh_IO.cpp:24-70) is built and traced on the GPU: the tree and
    the hits equal the compiled reference's recorded ones bit for bit, incoherent rays equal
    the oracle's, and a render equals the oracle's within 1e-4."""
    import os
    import torch
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_bunny_head.npz"))
    (tmp_path / "bunny_head.ply").write_bytes(g["ply"].tobytes())
    (tmp_path / "head.scn").write_bytes(g["scn"].tobytes())
    v, t = rtx.load_scene(tmp_path / "head.scn")
    assert np.array_equal(v.view(np.uint32), g["verts"].view(np.uint32)) and np.array_equal(t, g["tris"])
    b = rtx.build(v, t)
    nodes, tris, prims = b.export()
    assert_same_nodes(nodes, np.ascontiguousarray(g["nodes"]).view(O.NODE48).reshape(-1), "mesh DFS order")
    assert np.array_equal(tris, g["rtris"]) and np.array_equal(prims, g["prims"])
    for key in ("rays", "edge"):
        h = b.trace(g[key])
        ref = g[key + "_hit"]
        assert np.array_equal(np.stack([h["t"], h["u"], h["v"], h["tri"].view(np.float32)], 1).view(np.uint32),
                              ref.view(np.uint32)), key
    o = O.build(v, t)
    lo, hi = v[:, :3].min(0), v[:, :3].max(0)
    rays = S.incoherent_rays(1 << 16, seed=9)
    rays[:, :3] = lo + (hi - lo) * rays[:, :3]
    h, r = b.trace(rays), O.trace(o, rays, nthreads=16)
    assert np.array_equal(h["tri"], r["tri"]) and np.array_equal(h["t"].view(np.uint32), r["t"].view(np.uint32))
    assert (h["tri"] != 0xFFFFFFFF).mean() > 0.05
    c = (lo + hi) / 2
    cam = S.camera_lookat(c + np.array([0.0, 0.2, -1.2]) * (hi - lo).max(), c, 40.0)
    p = rtx.make_params(64, 48, 2, 4, 3, seed=3)
    img = rtx.render(b, cam, p).cpu().numpy()
    ref = O.render(o, cam, O.make_params(64, 48, 2, 4, 3, seed=3), nthreads=16)
    np.testing.assert_allclose(img, ref, rtol=1e-4, atol=1e-6)
    assert np.abs(img - img.mean()).max() > 0.05   # the mesh is in view, not only sky
    del torch
