"""GPU builder parity: Morton codes, sort order, BFS and DFS node arrays, reordered
triangles and original ids must equal the CPU oracle bit for bit."""
import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S
from scenes_util import assert_same_nodes, scene_suite

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def suite():
    return scene_suite(rtx.geo_sphere, rtx.geo_box)


def _check(name, v, t):
    print(f"[build parity] {name}: {len(t)} triangles", flush=True)
    g = rtx.build(v, t)
    o = O.build(v, t)
    codes, srt = g.morton()
    ocodes, _ = O.morton(v, t)
    assert np.array_equal(codes, ocodes), f"{name}: Morton codes differ"
    assert np.array_equal(srt, o.sorted), f"{name}: sort order differs"
    assert_same_nodes(g.export_bfs(), o.bfs_nodes, f"{name} (builder order)")
    nodes, tris, prims = g.export()
    assert_same_nodes(nodes, o.nodes, f"{name} (DFS order)")
    assert np.array_equal(tris, o.tris), f"{name}: reordered triangles differ"
    assert np.array_equal(prims, o.prims), f"{name}: original ids differ"
    inf = g.info()
    assert inf["tri_index_size"] == o.info["tindex_size"]
    assert inf["sorted_tris_size"] == o.info["sorted_size"]
    assert inf["max_depth"] >= o.info["depth"]
    return g, o


@pytest.mark.parametrize("name", ["rand1", "rand5", "rand17", "rand65", "rand100", "rand1000", "rand20000",
                                  "cornell", "test_traverse", "straddle", "dupcodes", "degenerate", "grid"])
def test_build_matches_oracle(suite, name):
    v, t = suite[name]
    _check(name, v, t)


def test_budget_exhaustion_path(suite):
    v, t = suite["straddle"]
    o = O.build(v, t)
    assert o.info["budget_rejected"] > 0, "scene no longer exercises the spatial-split budget"
    _check("straddle", v, t)


def test_budget_exhaustion_on_wide_levels():
    """Overlapping triangles in levels wider than one 1024-node chunk: the multi-block
    budget pass must leave levels whose spatial splits do not all fit to the in-order
    pass, and take the others itself."""
    v, t = S.random_triangles(30000, seed=8, mu=-1.0, sigma=0.3)
    o = O.build(v, t)
    assert o.info["budget_rejected"] > 0, "scene no longer exhausts the spatial-split budget"
    _check("straddle30k", v, t)


def test_build_1m_random():
    """The 1M-triangle benchmark scene at full size."""
    v, t = S.random_triangles(1_000_000, seed=1)
    _check("rand1M", v, t)


def test_build_rejects_bad_index():
    v, t = S.random_triangles(10, seed=3)
    t = t.copy()
    t[4, 1] = len(v) + 5
    with pytest.raises(rtx.RtxError, match="out-of-bounds"):
        rtx.build(v, t)


def test_build_device_inputs():
    import torch
    v, t = S.random_triangles(5000, seed=9)
    vd = torch.from_numpy(v).cuda()
    td = torch.from_numpy(t.view(np.int32)).cuda()
    g = rtx.build_device(vd, td)
    nodes, tris, prims = g.export()
    o = O.build(v, t)
    assert_same_nodes(nodes, o.nodes, "device inputs")
    assert np.array_equal(prims, o.prims)


def test_builds_reusing_cached_scratch(suite):
    """Build scratch is cached across builds, so later builds run on blocks that hold an
    earlier build's data: interleaved builds of different sizes (the larger one first, so
    smaller builds take its blocks) still equal the oracle bit for bit, and a handle
    traced after other builds gives the same hits as when it was new."""
    big = suite["rand20000"]
    g_big, _ = _check("rand20000 (first)", *big)
    rays = S.incoherent_rays(4096, seed=21)
    lo, hi = big[0][:, :3].min(0), big[0][:, :3].max(0)
    rays[:, :3] = lo + (hi - lo) * rays[:, :3]
    h0 = g_big.trace(rays)
    for name in ("rand17", "straddle", "rand1000", "dupcodes", "rand20000", "grid", "rand65"):
        _check(name + " (after others)", *suite[name])
    h1 = g_big.trace(rays)
    for k in ("t", "u", "v", "tri"):
        assert np.array_equal(np.asarray(h0[k]).view(np.uint32), np.asarray(h1[k]).view(np.uint32)), k
