"""The 4-wide traversal layout (tuning trace_tree 1) against the binary walk and the
compiled reference. The wide walk tests the same boxes of the same tree in another order,
so every ray's t, and its u and v whenever the triangle is the same, must be bit-identical;
the triangle index may differ only where two triangles give the same t bit for bit (with
spatial splits that is typically one triangle referenced from two leaves, which map to the
same original triangle). Any-hit answers must be identical."""
import glob
import os

import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S
from scenes_util import edge_rays, scene_suite

pytestmark = pytest.mark.gpu

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_*.npz"))
                if not p.endswith("reference_clip.npz"))


def _trace(g, rays, tree, **kw):
    import torch
    saved = rtx.get_tuning()
    try:
        rtx.set_tuning(trace_tree=tree, **kw)
        rd = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
        h = torch.zeros((len(rays), 4), device="cuda")
        occ = torch.zeros(len(rays), dtype=torch.uint8, device="cuda")
        g.trace_closest_device(rd, h)
        g.trace_any_device(rd, occ)
        torch.cuda.synchronize()
        return h.cpu().numpy(), occ.cpu().numpy()
    finally:
        rtx.set_tuning(**saved)


def assert_tie_equivalent(ref, got, prims, what):
    """ref, got: (N, 4) hits (t, u, v, tri bits). Returns (ties, same-original-triangle ties)."""
    rt, gt = ref[:, 3].view(np.uint32), got[:, 3].view(np.uint32)
    assert np.array_equal((rt == 0xFFFFFFFF), (gt == 0xFFFFFFFF)), f"{what}: hit/miss differs"
    hit = rt != 0xFFFFFFFF
    assert np.array_equal(ref[hit, 0].view(np.uint32), got[hit, 0].view(np.uint32)), \
        f"{what}: t differs on {(ref[hit, 0].view(np.uint32) != got[hit, 0].view(np.uint32)).sum()} rays"
    same = hit & (rt == gt)
    assert np.array_equal(ref[same, 1:3].view(np.uint32), got[same, 1:3].view(np.uint32)), f"{what}: u/v differ"
    tie = hit & (rt != gt)
    return int(tie.sum()), int((tie & (prims[np.minimum(rt, len(prims) - 1)] == prims[np.minimum(gt, len(prims) - 1)])).sum())


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[10:-4] for p in GOLDEN])
def test_wide_matches_reference_golden(path):
    g = np.load(path)
    b = rtx.build(g["verts"], g["tris"])
    _, _, prims = b.export()
    for key in ("rays", "edge"):
        for kw in (dict(trace_loop=0), dict(trace_loop=1, leaf_coop=1), dict(trace_loop=2, leaf_coop=0)):
            h, occ = _trace(b, g[key], 1, **kw)
            ties, dup = assert_tie_equivalent(g[key + "_hit"], h, prims, f"{os.path.basename(path)} {key} {kw}")
            assert np.array_equal(occ.astype(bool), g[key + "_hit"][:, 3].view(np.uint32) != 0xFFFFFFFF)
            print(f"[wide golden] {os.path.basename(path)} {key} {kw}: {ties} exact-t ties ({dup} same triangle)",
                  flush=True)


@pytest.mark.parametrize("name", ["grid", "degenerate", "cornell", "straddle", "dupcodes"])
def test_wide_edge_rays(name):
    v, t = scene_suite(rtx.geo_sphere, rtx.geo_box)[name]
    b = rtx.build(v, t)
    _, _, prims = b.export()
    rays = edge_rays(v, 80000, seed=9)
    ref, rocc = _trace(b, rays, 0)
    for kw in (dict(trace_loop=1), dict(trace_loop=2, ray_sort=1), dict(trace_loop=2, ray_sort=0, leaf_coop=0)):
        h, occ = _trace(b, rays, 1, **kw)
        ties, dup = assert_tie_equivalent(ref, h, prims, f"{name} {kw}")
        assert np.array_equal(occ, rocc)
        print(f"[wide edge] {name} {kw}: {ties} exact-t ties ({dup} same triangle)", flush=True)


def test_wide_full_headline_workload():
    """1M random triangles (seed 1), all 4,194,304 incoherent rays and the 2048^2 coherent
    camera rays: wide = binary up to exact-t ties."""
    v, t = S.random_triangles(1_000_000, seed=1)
    b = rtx.build(v, t)
    _, _, prims = b.export()
    cam = S.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 45.0)
    for name, rays in (("incoherent", S.incoherent_rays(4_194_304, seed=1)),
                       ("coherent", S.pinhole_rays(cam, 2048, 2048, tmin=0.0))):
        ref, rocc = _trace(b, rays, 0)
        h, occ = _trace(b, rays, 1)
        ties, dup = assert_tie_equivalent(ref, h, prims, name)
        assert np.array_equal(occ, rocc)
        assert ties <= 1e-4 * len(rays)
        print(f"[wide 1M] {name}: {ties} exact-t ties among {len(rays)} rays ({dup} the same original triangle)",
              flush=True)


def test_wide_on_loaded_and_tiny_trees(tmp_path):
    """Trees loaded from a .bvh file get the wide layout too; a tree whose root is a leaf has
    none and traces the binary way under either setting."""
    v, t = S.random_triangles(30_000, seed=8)
    b = rtx.build(v, t)
    b.save(tmp_path / "w.bvh")
    lb = rtx.load_bvh(tmp_path / "w.bvh")
    rays = S.incoherent_rays(50_000, seed=4)
    ref, _ = _trace(b, rays, 0)
    for g in (b, lb):
        h, _ = _trace(g, rays, 1)
        assert_tie_equivalent(ref, h, b.export()[2], "loaded")
    v, t = S.random_triangles(9, seed=2)
    small = rtx.build(v, t)
    assert small.info()["internal_count"] == 0
    rays = S.incoherent_rays(5000, seed=1)
    a, _ = _trace(small, rays, 0)
    c, _ = _trace(small, rays, 1)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    r = O.trace(O.build(v, t), rays)
    assert np.array_equal(a[:, 3].view(np.uint32), r["tri"])


def test_noalloc_entry_with_wide_tree():
    import ctypes as C
    import torch
    v, t = S.random_triangles(100_000, seed=3)
    b = rtx.build(v, t)
    rays = S.incoherent_rays(200_000, seed=6)
    ref, _ = _trace(b, rays, 1)
    saved = rtx.get_tuning()
    try:
        rtx.set_tuning(trace_tree=1)
        rd = torch.from_numpy(rays).cuda()
        h = torch.zeros((len(rays), 4), device="cuda")
        assert rtx.lib().rtx_trace_closest_noalloc(b.handle, C.c_void_p(rd.data_ptr()), len(rays),
                                                   C.c_void_p(h.data_ptr()), None, None) == 0
        torch.cuda.synchronize()
    finally:
        rtx.set_tuning(**saved)
    assert np.array_equal(h.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_wide_setting_on_a_tree_too_deep_for_its_stack(tmp_path):
    """A loaded caterpillar tree 60 levels deep (each internal node: a one-triangle leaf and
    the next internal node) is deeper than the wide walk's stack bound, so trace_tree 1 takes
    the binary walk: bit-identical hits, and the oracle's."""
    D = 60
    nodes = np.zeros(2 * D + 1, O.NODE48)
    tris = np.arange(3 * (D + 1), dtype=np.uint32).reshape(-1, 3)
    verts = np.zeros((3 * (D + 1), 6), np.float32)
    for k in range(D + 1):   # triangle k in the plane z = k, facing -z
        verts[3 * k:3 * k + 3, :3] = [[0, 0, k], [0, 1, k], [1, 0, k]]
        verts[3 * k:3 * k + 3, 5] = -1
    for i in range(D):
        nodes[2 * i]["is_leaf"], nodes[2 * i]["a"], nodes[2 * i]["b"] = 0, 2 * i + 1, 2 * i + 2
        nodes[2 * i]["bmin"][:3], nodes[2 * i]["bmax"][:3] = [0, 0, i], [1, 1, D]
        nodes[2 * i + 1]["is_leaf"], nodes[2 * i + 1]["a"], nodes[2 * i + 1]["b"] = 1, i, 1
        nodes[2 * i + 1]["bmin"][:3], nodes[2 * i + 1]["bmax"][:3] = [0, 0, i], [1, 1, i]
    nodes[2 * D]["is_leaf"], nodes[2 * D]["a"], nodes[2 * D]["b"] = 1, D, 1
    nodes[2 * D]["bmin"][:3], nodes[2 * D]["bmax"][:3] = [0, 0, D], [1, 1, D]
    hdr = np.array([1, len(nodes), len(tris), len(verts)], np.uint32)
    (tmp_path / "deep.bvh").write_bytes(hdr.tobytes() + nodes.tobytes() + tris.tobytes() + verts.tobytes())
    b = rtx.load_bvh(tmp_path / "deep.bvh")
    assert b.info()["max_depth"] == D
    rng = np.random.default_rng(3)
    rays = np.zeros((20000, 8), np.float32)
    rays[:, 0:2] = rng.random((20000, 2)) * 1.2 - 0.1
    rays[:, 2] = rng.uniform(-1, D + 1, 20000)
    d = rng.normal(size=(20000, 3)) * [0.05, 0.05, 1]
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf
    a, _ = _trace(b, rays, 0)
    w, _ = _trace(b, rays, 1)
    assert np.array_equal(a.view(np.uint32), w.view(np.uint32))
    ref = O.trace(O.Tree(nodes, tris, np.arange(len(tris), dtype=np.uint32), verts), rays)
    assert np.array_equal(a[:, 3].view(np.uint32), ref["tri"])
    assert (ref["tri"] != 0xFFFFFFFF).mean() > 0.2
