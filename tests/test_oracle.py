"""The CPU oracle against the compiled reference (oracle/_ref), library routines,
brute force and the reference's own known-answer tests."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from rtx import scenes as S
from scenes_util import assert_same_nodes, scene_suite

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")


@pytest.fixture(scope="module")
def suite():
    if O.ref_available():
        return scene_suite(O.ref_geo_sphere, O.ref_geo_box)
    import rtx
    return scene_suite(rtx.geo_sphere, rtx.geo_box)


@needs_ref
@pytest.mark.parametrize("name", ["rand1", "rand5", "rand17", "rand65", "rand100", "rand1000", "rand20000",
                                  "cornell", "test_traverse", "straddle", "dupcodes", "degenerate", "grid"])
def test_build_matches_compiled_reference(suite, name):
    v, t = suite[name]
    o = O.build(v, t)
    r = O.ref_build(v, t)
    codes, box = O.morton(v, t)
    assert np.array_equal(codes, r.codes)
    assert np.array_equal(box.view(np.uint32), r.scene_box.view(np.uint32))
    assert np.array_equal(o.sorted, r.sorted)
    assert_same_nodes(o.bfs_nodes, r.bfs_nodes, name + " builder order")
    assert_same_nodes(o.nodes, r.nodes, name)
    assert np.array_equal(o.tris, r.tris)
    assert np.array_equal(o.prims, r.prims)
    assert o.info["tindex_size"] == r.info["tindex_size"]
    assert o.info["sorted_size"] == r.info["sorted_size"]


def test_suite_exercises_all_builder_paths(suite):
    info = {k: O.build(*suite[k]).info for k in ("straddle", "rand20000", "dupcodes")}
    assert info["straddle"]["budget_rejected"] > 0
    assert info["rand20000"]["spatial"] > 0


def test_sort_is_stable_argsort():
    rng = np.random.default_rng(5)
    codes = rng.integers(0, 1 << 20, 50000, dtype=np.uint64) << np.uint64(30)
    codes[::7] = codes[3]
    assert np.array_equal(O.sort(codes), np.argsort(codes, kind="stable").astype(np.uint32))


def test_morton_known_answers():
    """Closed forms of the 21-bit interleave: a triangle at the scene-box minimum codes 0,
    at the maximum all 63 bits; x occupies bit 3i+2, y 3i+1, z 3i."""
    def tri_at(p):
        return np.array([[*p, 0, 0, 1]] * 3, np.float32)
    corners = [tri_at((0, 0, 0)), tri_at((1, 1, 1)), tri_at((1, 0, 0)), tri_at((0, 1, 0)), tri_at((0, 0, 1))]
    v = np.concatenate(corners)
    t = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    codes, _ = O.morton(v, t)
    full = (1 << 63) - 1
    assert codes[0] == 0 and codes[1] == full
    assert codes[2] == 0x4924924924924924 and codes[3] == 0x2492492492492492 and codes[4] == 0x1249249249249249


@needs_ref
def test_clip_matches_reference_known_answers():
    """clipTriangleToAABB cases of the reference's tests/test_aabb.cpp:527-594."""
    cases = [([1, 1, 1, 3, 1, 1, 2, 3, 1], [0, 0, 0, 10, 10, 10]),
             ([20, 20, 20, 30, 20, 20, 25, 30, 20], [0, 0, 0, 10, 10, 10]),
             ([-2, 0, 0, 2, 0, 0, 0, 2, 0], [0, -1, -1, 5, 5, 5]),
             ([-1, -1, 0, 5, -1, 0, -1, 5, 0], [0, 0, -1, 3, 3, 1]),
             ([0, 0, 0, 0, 4, 0, 3, 2, 0], [0, 0, -1, 5, 5, 1])]
    ok, b = O.ref_clip(*cases[0])
    assert ok and list(b) == [1, 1, 1, 3, 3, 1]
    ok, _ = O.ref_clip(*cases[1])
    assert not ok
    ok, b = O.ref_clip(*cases[2])
    assert ok and b[0] >= -1e-5 and b[3] <= 2 + 1e-5


def _brute(v, t, rays):
    """Nearest accepted triangle by brute force with the reference's Moller-Trumbore
    (bvh_traverse.cpp:47-87) in float32, vectorised over triangles."""
    P = v[:, :3][t.astype(np.int64)]
    v0, e1, e2 = P[:, 0], P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    out = np.full(len(rays), 0xFFFFFFFF, np.uint32)
    tout = np.zeros(len(rays), np.float32)
    f32 = np.float32

    def cross(a, b):
        return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

    def dot(a, b):
        return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + (a[..., 2] * b[..., 2] + f32(0))
    with np.errstate(all="ignore"):
        for i, r in enumerate(rays):
            o, d = r[0:3], r[4:7]
            h = cross(np.broadcast_to(d, e2.shape), e2)
            a = dot(e1, h)
            f = f32(1) / a
            s = o - v0
            u = f * dot(s, h)
            q = cross(s, e1)
            vv = f * dot(np.broadcast_to(d, q.shape), q)
            tt = f * dot(e2, q)
            ok = (a < 0) & (u >= 0) & (u <= 1) & (vv >= 0) & (u + vv <= 1) & (tt >= r[3]) & (tt <= r[7])
            if ok.any():
                tt = np.where(ok, tt, np.inf)
                k = int(np.argmin(tt))
                if np.count_nonzero(tt == tt[k]) == 1:
                    out[i], tout[i] = k, tt[k]
                else:
                    out[i], tout[i] = 0xFFFFFFFE, tt[k]   # exact tie: either id is valid
    return out, tout


@pytest.mark.parametrize("name", ["rand1000", "cornell", "test_traverse", "straddle"])
def test_trace_matches_brute_force(suite, name):
    v, t = suite[name]
    o = O.build(v, t)
    lo, hi = v[:, :3].min(0) - 0.1, v[:, :3].max(0) + 0.1
    rays = S.incoherent_rays(400, seed=3)
    rays[:, :3] = lo + (hi - lo) * rays[:, :3]
    h = O.trace(o, rays)
    bid, bt = _brute(v, t, rays)
    prim = np.where(h["tri"] != 0xFFFFFFFF, o.prims[np.minimum(h["tri"], len(o.prims) - 1)], 0xFFFFFFFF)
    sel = bid != 0xFFFFFFFE
    assert np.array_equal(prim[sel], bid[sel])
    hit = bid != 0xFFFFFFFF
    assert np.array_equal(h["t"][hit].view(np.uint32), bt[hit].view(np.uint32))


@needs_ref
@pytest.mark.parametrize("name", ["rand20000", "cornell", "test_traverse", "straddle", "dupcodes", "degenerate",
                                  "grid"])
def test_trace_matches_compiled_reference(suite, name):
    v, t = suite[name]
    r = O.ref_build(v, t)
    lo, hi = v[:, :3].min(0) - 0.1, v[:, :3].max(0) + 0.1
    rays = S.incoherent_rays(20000, seed=4)
    rays[:, :3] = lo + (hi - lo) * rays[:, :3]
    a = O.trace(r, rays, normals=True, any_hit=False)
    b = O.ref_trace(r, rays, normals=True)
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert np.array_equal(a["tri"], b["tri"])
    assert np.array_equal(a["normal"].view(np.uint32), b["normal"].view(np.uint32))
    anyh = O.trace(r, rays, any_hit=True)
    assert np.array_equal(anyh["tri"] != 0xFFFFFFFF, b["tri"] != 0xFFFFFFFF)


@needs_ref
@pytest.mark.parametrize("name", ["grid", "degenerate", "cornell", "rand1000"])
def test_edge_rays_match_compiled_reference(suite, name):
    """Zero, -0 and denormal direction components, origins on box planes (0 * inf in the
    slab test) and tmin/tmax windows: the oracle traversal equals the reference bit for bit."""
    from scenes_util import edge_rays
    v, t = suite[name]
    r = O.ref_build(v, t)
    rays = edge_rays(v, 8000, seed=5)
    a = O.trace(r, rays, normals=True)
    b = O.ref_trace(r, rays, normals=True)
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert np.array_equal(a["tri"], b["tri"])
    assert np.array_equal(a["normal"].view(np.uint32), b["normal"].view(np.uint32))
    if name == "grid":
        assert (a["tri"] != 0xFFFFFFFF).mean() > 0.2


def test_reference_traverse_known_answers(suite):
    """tests/test_traverse.cpp:117-236 rays on its scene, through the oracle."""
    v, t = suite["test_traverse"]
    o = O.build(v, t)

    def ray(p, d):
        d = np.asarray(d, np.float32)
        return np.array([*p, 0.0, *(d / np.float32(np.linalg.norm(d))), np.inf], np.float32)
    rays = np.stack([ray((0, 0, 0), (0, 0, 1)), ray((0, 0, 8), (0, 0, 1)), ray((0, 0, 15), (0, 0, -1)),
                     ray((-5, 0, 5), (1, 0, 0)), ray((0, 0, 8), (0, 0, -1)), ray((3, 0, 0), (0, 0, 1)),
                     ray((0, 10, 0), (0, 0, 1)), ray((0, 0, 0), (0, 0, -1)), ray((0, 0, 0), (0, 1, 0)),
                     ray((3, 0, 0), (0, 1, 0)), ray((0, -5, 0), (0, 1, 0)), ray((25, 0, 0), (0, -1, 0))])
    h = O.trace(o, rays, normals=True)
    hit = list(h["tri"] != 0xFFFFFFFF)
    assert hit == [True] * 5 + [False] * 3 + [True, False, True, False]
    assert 3.9 <= h["t"][0] <= 4.1 and h["normal"][0, 2] < 0
    assert 2.5 <= h["t"][8] <= 3.5 and h["normal"][8, 1] < 0
    assert 1.9 <= h["t"][10] <= 2.2 and h["normal"][10, 1] > 0


def test_sincos_and_sample_table():
    xs = np.linspace(0, 2 * math.pi, 2001, dtype=np.float32)
    err = max(max(abs(O.sincos(x)[0] - math.sin(float(x))), abs(O.sincos(x)[1] - math.cos(float(x)))) for x in xs)
    assert err < 3e-7
    u = np.array([O.sample(3, p, s, d) for p in range(40) for s in range(10) for d in range(8)])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.02
    assert O.sample(1, 2, 3, 4) != O.sample(1, 2, 3, 5)


def test_atan2_spec_against_numpy():
    """The IBL lookup's atan2 (octant reduction + polynomial) against numpy.arctan2 over
    all directions of a dense angle grid and the axis/diagonal special cases: within
    2e-6 rad, exact quadrant signs, and exact 0, +-pi/2, pi on the axes."""
    ang = np.linspace(-np.pi, np.pi, 20001)
    err = 0.0
    for r in (1e-3, 1.0, 7.5):
        for a in ang:
            y, x = np.float32(r * np.sin(a)), np.float32(r * np.cos(a))
            got = O.atan2(y, x)
            err = max(err, abs(got - float(np.arctan2(np.float64(y), np.float64(x)))))
    assert err < 2e-6, err
    assert O.atan2(0.0, 1.0) == 0.0
    assert abs(O.atan2(1.0, 0.0) - np.pi / 2) < 1e-7 and abs(O.atan2(-1.0, 0.0) + np.pi / 2) < 1e-7
    assert abs(O.atan2(0.0, -1.0) - np.pi) < 1e-6
    assert abs(O.atan2(1.0, 1.0) - np.pi / 4) < 1e-6 and abs(O.atan2(-1.0, -1.0) + 3 * np.pi / 4) < 1e-6
