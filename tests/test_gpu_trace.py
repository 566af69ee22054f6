"""GPU traversal parity: same hit triangle (reference triIndex), bit-identical t/u/v,
any-hit == closest-hit.hit, identical per-ray work counts, on the GPU-built tree."""
import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S
from scenes_util import scene_suite

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def suite():
    return scene_suite(rtx.geo_sphere, rtx.geo_box)


def _rays_for(v, n, seed):
    lo = v[:, :3].min(0) - 0.1
    hi = v[:, :3].max(0) + 0.1
    r = S.incoherent_rays(n, seed=seed)
    r[:, 0:3] = lo + (hi - lo) * r[:, 0:3]
    return r


def _device_trace(g, rays):
    import torch
    rd = torch.from_numpy(rays).cuda()
    hits = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    nrm = torch.zeros((len(rays), 3), dtype=torch.float32, device="cuda")
    occ = torch.zeros(len(rays), dtype=torch.uint8, device="cuda")
    cnt = torch.zeros((len(rays), 2), dtype=torch.int32, device="cuda")
    g.trace_closest_device(rd, hits, nrm)
    g.trace_any_device(rd, occ)
    g.trace_counts_device(rd, cnt)
    torch.cuda.synchronize()
    h = hits.cpu().numpy()
    return h, nrm.cpu().numpy(), occ.cpu().numpy(), cnt.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("name", ["rand1", "rand17", "rand1000", "rand20000", "cornell", "test_traverse",
                                  "straddle", "dupcodes", "degenerate", "grid"])
def test_trace_matches_oracle(suite, name):
    v, t = suite[name]
    print(f"[trace parity] {name}", flush=True)
    g = rtx.build(v, t)
    o = O.build(v, t)
    rays = _rays_for(v, 20000, seed=len(t))
    h, nrm, occ, cnt = _device_trace(g, rays)
    ref = O.trace(o, rays, counts=True, normals=True, nthreads=8)
    tri = h[:, 3].view(np.uint32)
    assert np.array_equal(tri, ref["tri"]), f"{name}: hit ids differ on {(tri != ref['tri']).sum()} rays"
    for k, col in (("t", 0), ("u", 1), ("v", 2)):
        assert np.array_equal(h[:, col].view(np.uint32), ref[k].view(np.uint32)), f"{name}: {k} differs"
    assert np.array_equal(occ.astype(bool), ref["tri"] != 0xFFFFFFFF), f"{name}: any-hit != closest hit"
    assert np.array_equal(cnt, ref["counts"]), f"{name}: per-ray work counts differ"
    np.testing.assert_allclose(nrm, ref["normal"], rtol=0, atol=1e-6)
    if name not in ("dupcodes", "rand1"):
        assert (tri != 0xFFFFFFFF).any()


@pytest.mark.parametrize("name", ["grid", "degenerate", "cornell"])
def test_edge_rays_all_variants(suite, name):
    """Zero, -0 and denormal direction components and origins on box planes (the exact
    NaN-aware slab test) mixed with ordinary rays (the fast slab test) in the same waves:
    every loop shape, with and without ray sorting, equals the oracle bit for bit, and so
    do the counts."""
    import torch
    from scenes_util import edge_rays
    v, t = suite[name]
    g = rtx.build(v, t)
    o = O.build(v, t)
    rays = edge_rays(v, 80000, seed=9)
    ref = O.trace(o, rays, counts=True, normals=True, nthreads=8)
    saved = rtx.get_tuning()
    try:
        for loop, sort, wait, coop in ((0, 0, 0, 0), (1, 0, 8, 0), (2, 1, 16, 0), (2, 0, 0, 0), (1, 0, 32, 1),
                                       (2, 1, 16, 1)):
            rtx.set_tuning(trace_loop=loop, ray_sort=sort, wait_lanes=wait, leaf_coop=coop)
            print(f"[edge rays] {name} loop={loop} sort={sort} coop={coop}", flush=True)
            h, nrm, occ, cnt = _device_trace(g, rays)
            assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"]), (loop, sort)
            for k, c in (("t", 0), ("u", 1), ("v", 2)):
                hit = ref["tri"] != 0xFFFFFFFF
                assert np.array_equal(h[hit, c].view(np.uint32), ref[k][hit].view(np.uint32)), (k, loop, sort)
            assert np.array_equal(occ.astype(bool), ref["tri"] != 0xFFFFFFFF), (loop, sort)
            assert np.array_equal(cnt, ref["counts"]), (loop, sort)
    finally:
        rtx.set_tuning(**saved)


def test_trace_host_wrapper_and_reference_rays(suite):
    """The 12 rays of the reference's tests/test_traverse.cpp:117-236 through the GPU."""
    v, t = suite["test_traverse"]
    g = rtx.build(v, t)

    def ray(o, d):
        d = np.asarray(d, np.float32)
        d = d / np.float32(np.sqrt(np.float32((d * d).sum())))
        return np.array([*o, 0.0, *d, np.inf], np.float32)
    rays = np.stack([ray((0, 0, 0), (0, 0, 1)), ray((0, 0, 8), (0, 0, 1)), ray((0, 0, 15), (0, 0, -1)),
                     ray((-5, 0, 5), (1, 0, 0)), ray((0, 0, 8), (0, 0, -1)), ray((3, 0, 0), (0, 0, 1)),
                     ray((0, 10, 0), (0, 0, 1)), ray((0, 0, 0), (0, 0, -1)), ray((0, 0, 0), (0, 1, 0)),
                     ray((3, 0, 0), (0, 1, 0)), ray((0, -5, 0), (0, 1, 0)), ray((25, 0, 0), (0, -1, 0))])
    h = g.trace(rays, normals=True)
    hit = h["tri"] != 0xFFFFFFFF
    assert list(hit) == [True] * 5 + [False] * 3 + [True, False, True, False]
    t_ = h["t"]
    assert 3.9 <= t_[0] <= 4.1 and h["normal"][0, 2] < 0
    assert 2.9 <= t_[1] <= 3.1 and h["normal"][1, 2] < 0
    assert 1.9 <= t_[2] <= 2.1 and h["normal"][2, 2] > 0
    assert 3.9 <= t_[3] <= 4.1 and h["normal"][3, 0] < 0
    assert 1.9 <= t_[4] <= 2.1 and h["normal"][4, 2] > 0
    assert 2.5 <= t_[8] <= 3.5 and h["normal"][8, 1] < 0
    assert 1.9 <= t_[10] <= 2.2 and h["normal"][10, 1] > 0


def test_trace_1m_subset():
    """1M-triangle scene: a 65k-ray subset of both benchmark ray sets against the oracle."""
    v, t = S.random_triangles(1_000_000, seed=1)
    g = rtx.build(v, t)
    o = O.build(v, t)
    inc = S.incoherent_rays(1 << 16, seed=1)
    cam = S.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 45.0)
    coh = S.pinhole_rays(cam, 2048, 2048)[:: 64]
    for name, rays in (("incoherent", inc), ("coherent", coh)):
        print(f"[trace 1M] {name}", flush=True)
        h, _, occ, cnt = _device_trace(g, rays)
        ref = O.trace(o, rays, counts=True, nthreads=16)
        assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"]), name
        assert np.array_equal(h[:, 0].view(np.uint32), ref["t"].view(np.uint32)), name
        assert np.array_equal(occ.astype(bool), ref["tri"] != 0xFFFFFFFF), name
        assert np.array_equal(cnt, ref["counts"]), name


@pytest.mark.parametrize("name", ["rand1", "rand1000", "rand20000", "cornell", "straddle", "dupcodes", "degenerate",
                                  "grid"])
def test_cooperative_leaves_match_oracle(suite, name):
    """Cooperative leaf phases (the pairs of all lanes at leaves dealt over the wave, a
    segment running over several 64-pair rounds when a leaf is large) give the oracle's
    hits, t/u/v bits and work counts, in the while-while and the persistent kernel."""
    v, t = suite[name]
    g = rtx.build(v, t)
    o = O.build(v, t)
    rays = _rays_for(v, 20000, seed=len(t) + 1)
    ref = O.trace(o, rays, counts=True, normals=True, nthreads=8)
    saved = rtx.get_tuning()
    try:
        for loop, wait in ((1, 0), (1, 32), (2, 16), (2, 64)):
            rtx.set_tuning(trace_loop=loop, wait_lanes=wait, leaf_coop=1)
            print(f"[coop parity] {name} loop={loop} wait={wait}", flush=True)
            h, nrm, occ, cnt = _device_trace(g, rays)
            assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"]), (loop, wait)
            for k, col in (("t", 0), ("u", 1), ("v", 2)):
                assert np.array_equal(h[:, col].view(np.uint32), ref[k].view(np.uint32)), (k, loop, wait)
            assert np.array_equal(occ.astype(bool), ref["tri"] != 0xFFFFFFFF), (loop, wait)
            assert np.array_equal(cnt, ref["counts"]), (loop, wait)
            np.testing.assert_allclose(nrm, ref["normal"], rtol=0, atol=1e-6)
    finally:
        rtx.set_tuning(**saved)


def test_tuning_variants_give_identical_hits():
    """Every launch-shape variant (LDS stack depth, loop structure, ray sorting, how many
    lanes the leaf phase leaves behind) returns the same hits at each ray's own index."""
    import torch
    v, t = S.random_triangles(200_000, seed=5)
    g = rtx.build(v, t)
    rays = torch.from_numpy(S.incoherent_rays(1 << 16, seed=6)).cuda()
    hits = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    saved = rtx.get_tuning()
    ref = None
    try:
        combos = [(sort, depth, loop, 0, 0) for sort in (0, 1) for depth in (8, 12, 16) for loop in (0, 1, 2)]
        combos += [(sort, 8, loop, wait, 0) for sort in (0, 1) for loop in (1, 2) for wait in (1, 8, 32, 64)]
        combos += [(sort, depth, loop, wait, 1) for sort in (0, 1) for depth in (8, 16) for loop in (1, 2)
                   for wait in (0, 16, 64)]
        for sort, depth, loop, wait, coop in combos:
            rtx.set_tuning(trace_lds_depth=depth, trace_loop=loop, ray_sort=sort, wait_lanes=wait, leaf_coop=coop)
            hits.fill_(-1.0)
            g.trace_closest_device(rays, hits)
            torch.cuda.synchronize()
            h = hits.view(torch.int32).cpu().numpy().copy()
            if ref is None:
                ref = h
            assert np.array_equal(h, ref), f"variant depth={depth} loop={loop} sort={sort} wait={wait} coop={coop} differs"
    finally:
        rtx.set_tuning(**saved)


def test_ray_sort_counts_any_hit_and_normals(suite):
    """With ray sorting on, closest hit, any-hit, counts and normals still land at each
    ray's own index (70001 rays: above the sorting threshold, ragged tail)."""
    saved = rtx.get_tuning()
    try:
        rtx.set_tuning(ray_sort=1)
        for name in ("rand20000", "straddle", "cornell"):
            v, t = suite[name]
            g = rtx.build(v, t)
            o = O.build(v, t)
            rays = _rays_for(v, 70001, seed=11)
            print(f"[ray_sort] {name}", flush=True)
            h, nrm, occ, cnt = _device_trace(g, rays)
            ref = O.trace(o, rays, counts=True, normals=True, nthreads=8)
            assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"]), name
            assert np.array_equal(h[:, 0].view(np.uint32), ref["t"].view(np.uint32)), name
            assert np.array_equal(occ.astype(bool), ref["tri"] != 0xFFFFFFFF), name
            assert np.array_equal(cnt, ref["counts"]), name
            np.testing.assert_allclose(nrm, ref["normal"], atol=1e-6)
    finally:
        rtx.set_tuning(**saved)


def test_persistent_variant_counts_and_any_hit(suite):
    """The persistent dynamic-fetch kernel serves closest hit, counts, normals and any-hit
    exactly while lanes take many rays each (grid forced to 16 blocks). The camera rays
    mostly miss the root box, so whole groups of lanes go idle the moment they start a
    ray; the kernel's refill and exit must stay wave-uniform for them, for refill
    thresholds from 1 idle lane to all 64."""
    saved = rtx.get_tuning()
    try:
        for name, sort, coop, refill in (("rand20000", 0, 0, 18), ("straddle", 0, 0, 1), ("cornell", 0, 0, 64),
                                         ("rand20000", 1, 0, 32), ("rand20000", 0, 1, 64), ("straddle", 1, 1, 1),
                                         ("cornell", 0, 1, 18), ("cornell", 1, 1, 1)):
            rtx.set_tuning(trace_loop=2, trace_lds_depth=8, trace_grid_blocks=16, ray_sort=sort, wait_lanes=8 * sort,
                           leaf_coop=coop, refill_lanes=refill)
            v, t = suite[name]
            g = rtx.build(v, t)
            o = O.build(v, t)
            lo, hi = v[:, :3].min(0), v[:, :3].max(0)
            c = (lo + hi) / 2
            span = float((hi - lo).max())
            cam = S.camera_lookat(tuple(c - np.array([0.0, 0.0, 6.0 * span])), tuple(c), 60.0)
            camrays = S.pinhole_rays(cam, 256, 256)
            for kind, rays in (("scattered", _rays_for(v, 30001, seed=11)), ("camera", camrays)):
                print(f"[persistent] {name} sort={sort} coop={coop} refill={refill} {kind}", flush=True)
                h, nrm, occ, cnt = _device_trace(g, rays)
                ref = O.trace(o, rays, counts=True, normals=True, nthreads=8)
                if kind == "camera":
                    assert (ref["tri"] == 0xFFFFFFFF).mean() > 0.5
                assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"]), (name, kind)
                assert np.array_equal(h[:, 0].view(np.uint32), ref["t"].view(np.uint32)), (name, kind)
                assert np.array_equal(occ.astype(bool), ref["tri"] != 0xFFFFFFFF), (name, kind)
                assert np.array_equal(cnt, ref["counts"]), (name, kind)
                np.testing.assert_allclose(nrm, ref["normal"], atol=1e-6)
    finally:
        rtx.set_tuning(**saved)


def test_kernel_timer_counts_traversal_launches():
    """rtx_profile_begin/end time exactly the traversal kernels of the trace calls made in
    between (sorting excluded, so never more than the whole calls)."""
    import torch
    v, t = S.random_triangles(50_000, seed=3)
    g = rtx.build(v, t)
    rays = torch.from_numpy(S.incoherent_rays(1 << 17, seed=4)).cuda()
    hits = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    occ = torch.zeros(len(rays), dtype=torch.uint8, device="cuda")
    saved = rtx.get_tuning()
    try:
        for loop in (1, 2):
            rtx.set_tuning(trace_loop=loop, ray_sort=1)
            g.trace_closest_device(rays, hits)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            with rtx.KernelTimer() as kt:
                a.record()
                for _ in range(3):
                    g.trace_closest_device(rays, hits)
                g.trace_any_device(rays, occ)
                b.record()
            torch.cuda.synchronize()
            assert kt.launches == 4
            assert 0.0 < kt.ms <= a.elapsed_time(b) * 1.01 + 0.05
        with rtx.KernelTimer() as kt:
            pass
        assert kt.launches == 0 and kt.ms == 0.0
    finally:
        rtx.set_tuning(**saved)


def test_concurrent_streams_share_one_handle():
    """Trace calls on two streams against one handle (persistent kernel + ray sorting, both
    of which use the handle's scratch) give the same hits as the calls made one at a time.
    Both ray sets have the same length."""
    import torch
    v, t = S.random_triangles(100_000, seed=8)
    g = rtx.build(v, t)
    n = 1 << 17
    ra = torch.from_numpy(S.incoherent_rays(n, seed=21)).cuda()
    rb = torch.from_numpy(S.incoherent_rays(n, seed=22)).cuda()
    saved = rtx.get_tuning()
    try:
        rtx.set_tuning(trace_loop=2, ray_sort=1)
        ha = torch.zeros((n, 4), device="cuda")
        hb = torch.zeros((n, 4), device="cuda")
        g.trace_closest_device(ra, ha)
        g.trace_closest_device(rb, hb)
        torch.cuda.synchronize()
        ref_a, ref_b = ha.view(torch.int32).cpu().clone(), hb.view(torch.int32).cpu().clone()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        for rep in range(4):
            ha.fill_(-1.0)
            hb.fill_(-1.0)
            torch.cuda.synchronize()
            with torch.cuda.stream(s1):
                g.trace_closest_device(ra, ha, stream=s1)
            with torch.cuda.stream(s2):
                g.trace_closest_device(rb, hb, stream=s2)
            torch.cuda.synchronize()
            assert torch.equal(ha.view(torch.int32).cpu(), ref_a), rep
            assert torch.equal(hb.view(torch.int32).cpu(), ref_b), rep
    finally:
        rtx.set_tuning(**saved)


def test_reciprocal_matches_ieee_division_on_all_floats():
    """The traversal computes 1/a (Moller-Trumbore, bvh_traverse.cpp:62) and 1/dir
    (rayAABB, :26-28) as v_rcp_f32 plus one FMA Newton step where 2^-126 <= |x| < 2^126.
    Every one of the 2^32 float inputs must give the bits of the IEEE quotient 1.0f / x."""
    bad, first = rtx.debug_recip_check(0, 1 << 32)
    assert bad == 0, f"{bad} inputs differ from 1.0f / x, first 0x{first:08x}"
    # sub-ranges and the argument checks
    assert rtx.debug_recip_check(0x3F800000, 1 << 23) == (0, None)
    assert rtx.debug_recip_check(0xFFFFFFFF, 1) == (0, None)
    with pytest.raises(rtx.RtxError):
        rtx.debug_recip_check(0xFFFFFFFF, 2)


def test_empty_inputs():
    """No triangles is an error (the reference's bake has no defined result for it); zero
    rays is a no-op on every trace entry point, null buffers included."""
    import torch
    with pytest.raises(rtx.RtxError, match="no triangles"):
        rtx.build(np.zeros((0, 6), np.float32), np.zeros((0, 3), np.uint32))
    v, t = S.random_triangles(100, seed=2)
    g = rtx.build(v, t)
    out = g.trace(np.zeros((0, 8), np.float32), normals=True)
    assert len(out["t"]) == 0 and out["normal"].shape == (0, 3)
    rd = torch.zeros((0, 8), device="cuda")
    g.trace_closest_device(rd, torch.zeros((0, 4), device="cuda"))
    g.trace_any_device(rd, torch.zeros(0, dtype=torch.uint8, device="cuda"))
    g.trace_counts_device(rd, torch.zeros((0, 2), dtype=torch.int32, device="cuda"))
    torch.cuda.synchronize()


def test_device_buffers_are_checked():
    """The Python layer refuses buffers the kernels would index out of bounds or misread:
    short outputs, wrong dtypes, non-contiguous views, host tensors, short material arrays."""
    import torch
    v, t = S.random_triangles(500, seed=4)
    g = rtx.build(v, t)
    rd = torch.from_numpy(_rays_for(v, 100, seed=1)).cuda()
    bad = [lambda: g.trace_closest_device(rd, torch.zeros((99, 4), device="cuda")),
           lambda: g.trace_closest_device(rd, torch.zeros((100, 4), dtype=torch.float64, device="cuda")),
           lambda: g.trace_closest_device(rd, torch.zeros((4, 100), device="cuda").t()),
           lambda: g.trace_closest_device(rd.cpu(), torch.zeros((100, 4), device="cuda")),
           lambda: g.trace_closest_device(rd, torch.zeros((100, 4), device="cuda"), torch.zeros((100, 2), device="cuda")),
           lambda: g.trace_any_device(rd, torch.zeros(50, dtype=torch.uint8, device="cuda")),
           lambda: g.trace_counts_device(rd, torch.zeros((100, 2), device="cuda")),
           lambda: g.trace_closest_device(rd[:, :4].contiguous(), torch.zeros((100, 4), device="cuda"))]
    for k, f in enumerate(bad):
        with pytest.raises(ValueError):
            f()
    cam = S.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 45.0)
    p = rtx.make_params(8, 8, 1, 2, 3, seed=1)
    with pytest.raises(ValueError):
        rtx.render(g, cam, p, torch.zeros((499, 3), device="cuda"))
    with pytest.raises(ValueError):
        rtx.render(g, cam, p, torch.zeros((500, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        rtx.render(g, cam, p, None, None, out=torch.zeros((8, 7, 3), device="cuda"))
    img = rtx.render(g, cam, p, torch.full((500, 3), 0.5, device="cuda"))   # correct buffers pass
    torch.cuda.synchronize()
    assert img.shape == (8, 8, 3)


def test_host_trace_small_and_large_calls():
    """rtx_trace_closest_host: calls of up to 4096 rays run on pinned host buffers that the
    kernel reads and writes directly, larger ones through device staging; both give the
    device entry point's results, with and without normals, one ray at a time too."""
    import torch
    v, t = S.random_triangles(20000, seed=12)
    g = rtx.build(v, t)
    for n in (1, 7, 4096, 4097, 70000):
        rays = _rays_for(v, n, seed=n)
        h = g.trace(rays, normals=True)
        rd = torch.from_numpy(rays).cuda()
        hd = torch.zeros((n, 4), device="cuda")
        nd = torch.zeros((n, 3), device="cuda")
        g.trace_closest_device(rd, hd, nd)
        torch.cuda.synchronize()
        ref = hd.cpu().numpy()
        got = np.stack([h["t"], h["u"], h["v"], h["tri"].view(np.float32)], 1)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), n
        assert np.array_equal(h["normal"].view(np.uint32), nd.cpu().numpy().view(np.uint32)), n
        h2 = g.trace(rays)
        assert np.array_equal(h2["tri"], h["tri"]), n
    one = _rays_for(v, 200, seed=3)
    hits = [g.trace(one[i:i + 1])["tri"][0] for i in range(200)]
    assert np.array_equal(np.array(hits, np.uint32), g.trace(one)["tri"])


def _chunked_trace(g, rays, chunk):
    """Closest hits, normals and any-hit flags of rays traced in calls of `chunk` rays (calls
    of at most 16,384 rays take the one-wave-per-ray kernel)."""
    import torch
    rd = torch.from_numpy(rays).cuda()
    hits = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    nrm = torch.zeros((len(rays), 3), dtype=torch.float32, device="cuda")
    occ = torch.zeros(len(rays), dtype=torch.uint8, device="cuda")
    for a in range(0, len(rays), chunk):
        b = min(a + chunk, len(rays))
        g.trace_closest_device(rd[a:b], hits[a:b], nrm[a:b])
        g.trace_any_device(rd[a:b], occ[a:b])
    torch.cuda.synchronize()
    return hits.cpu().numpy(), nrm.cpu().numpy(), occ.cpu().numpy()


@pytest.mark.parametrize("name", ["rand1", "rand17", "rand20000", "cornell", "test_traverse", "straddle", "dupcodes",
                                  "degenerate", "grid"])
def test_wave_per_ray_kernel_matches_oracle(suite, name):
    """Calls of at most 16,384 rays walk each ray's tree breadth-first with one wave and keep
    the reference's tie order by path keys: ids, t, u, v and normals equal the oracle's bit for
    bit on random and edge-case rays (the grid scene's rays hit shared edges and vertices,
    the straddle scene's duplicated triangles), in calls of 256, 37 and 1 rays."""
    from scenes_util import edge_rays
    v, t = suite[name]
    g = rtx.build(v, t)
    o = O.build(v, t)
    rays = np.concatenate([_rays_for(v, 3000, seed=len(t) + 1), edge_rays(v, 3000, seed=5)])
    ref = O.trace(o, rays, normals=True, nthreads=8)
    hit = ref["tri"] != 0xFFFFFFFF
    for chunk in (256, 37, 1):
        print(f"[wave kernel] {name} chunk={chunk}", flush=True)
        sel = slice(0, 300) if chunk == 1 else slice(0, len(rays))
        h, nrm, occ = _chunked_trace(g, rays[sel], chunk)
        r_tri, r_hit = ref["tri"][sel], hit[sel]
        assert np.array_equal(h[:, 3].view(np.uint32), r_tri), f"{name}/{chunk}: ids differ on {(h[:, 3].view(np.uint32) != r_tri).sum()} rays"
        for k, c in (("t", 0), ("u", 1), ("v", 2)):
            assert np.array_equal(h[r_hit, c].view(np.uint32), ref[k][sel][r_hit].view(np.uint32)), (name, chunk, k)
        assert np.array_equal(occ.astype(bool), r_hit), (name, chunk)
        np.testing.assert_allclose(nrm, ref["normal"][sel], rtol=0, atol=1e-6)


def test_wave_per_ray_kernel_1m_and_ties():
    """The 1M-triangle scene: 16,384 incoherent and 4,096 coherent rays in calls of 256 rays
    equal the oracle bit for bit, and equal one large call of the same rays (the persistent
    kernel)."""
    v, t = S.random_triangles(1_000_000, seed=1)
    g = rtx.build(v, t)
    o = O.build(v, t)
    cam = S.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 45.0)
    rays = np.concatenate([S.incoherent_rays(16384, seed=7), S.pinhole_rays(cam, 2048, 2048)[::1024]])
    ref = O.trace(o, rays, nthreads=16)
    h, _, occ = _chunked_trace(g, rays, 256)
    big, _, occ_big = _chunked_trace(g, rays, len(rays))
    assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"])
    assert np.array_equal(h[:, 0].view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(h.view(np.uint32), big.view(np.uint32))
    assert np.array_equal(occ, occ_big)


def test_wave_per_ray_kernel_falls_back_on_wide_levels():
    """A ray along a stack of 40,000 parallel squares meets every node's box: its
    breadth-first walk reaches 1,024 internal nodes at one level (counted on the oracle's
    tree) before any leaf, more than the 512-entry lists hold, so the ray takes the one-lane walk:
    the hits still equal the oracle's (rays along the stack and ordinary ones, in small and
    large calls)."""
    n = 40000
    z = np.linspace(0.0, 1.0, n, dtype=np.float32)
    quad = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    verts = np.zeros((4 * n, 6), np.float32)
    verts[:, 0:2] = np.tile(quad, (n, 1))
    verts[:, 2] = np.repeat(z, 4)
    verts[:, 5] = -1.0
    base = 4 * np.arange(n, dtype=np.uint32)
    tris = np.concatenate([np.stack([base, base + 2, base + 1], 1), np.stack([base, base + 3, base + 2], 1)])
    g = rtx.build(verts, tris)
    o = O.build(verts, tris)
    rng = np.random.default_rng(3)
    along = np.zeros((64, 8), np.float32)
    along[:, 0:2] = rng.uniform(0.05, 0.95, (64, 2))
    along[:, 2] = 1.5
    along[:, 6] = -1.0          # straight down the stack: every square is hit
    along[:, 7] = np.inf
    rays = np.concatenate([along, _rays_for(verts, 1000, seed=4)])
    ref = O.trace(o, rays, nthreads=8)
    assert (ref["tri"][:64] != 0xFFFFFFFF).all()
    for chunk in (len(rays), 64, 1):
        print(f"[wave fallback] chunk={chunk}", flush=True)
        sel = slice(0, 128) if chunk == 1 else slice(0, len(rays))
        h, _, occ = _chunked_trace(g, rays[sel], chunk)
        assert np.array_equal(h[:, 3].view(np.uint32), ref["tri"][sel]), chunk
        hit = ref["tri"][sel] != 0xFFFFFFFF
        assert np.array_equal(h[hit, 0].view(np.uint32), ref["t"][sel][hit].view(np.uint32)), chunk
        assert np.array_equal(occ.astype(bool), hit), chunk
