"""GPU parity on the benchmark workloads at their real scene sizes:

* the 8,380,418-triangle terrain of the inverse render (2048^2 fBm heightfield, seed 4):
  the GPU tree equals the oracle's bit for bit (Morton codes, sort order, builder-order and
  DFS-order nodes, reordered triangles, original ids, index-array sizes), and a reduced
  render with texture albedo under sun + sky equals the oracle's within 1e-4;
* the 64 scenes of the batched render (seed 2): every tree equals the oracle's, reduced
  renders of the smallest, median and largest scene equal the oracle's, and at full size
  the graph-batched and shared-launch renders equal the plain one bit for bit;
* the differentiable render at 512 x 512: adjoint gradients of 64 albedo parameters drawn
  uniformly (and of the emitter) against central finite differences of the oracle.

Reference: bvh_bfs.cpp:72-134 and bvh_sah.cpp:390-394 (FIFO build and the spatial-split
budget on wide levels), main.cpp:145-204 (the path-trace loop)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S
from scenes_util import assert_same_nodes

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _tensor(a):
    import torch
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _close(gpu, orc, what):
    bit = np.mean(gpu.view(np.uint32) == orc.view(np.uint32))
    print(f"[render parity] {what}: {bit * 100:.3f}% of pixel channels bit-identical", flush=True)
    np.testing.assert_allclose(gpu, orc, rtol=RTOL, atol=1e-6, err_msg=what)


def _same_tree(g, o, what):
    nodes, tris, prims = g.export()
    assert_same_nodes(nodes, o.nodes, what + " DFS order")
    assert np.array_equal(tris, o.tris), what
    assert np.array_equal(prims, o.prims), what
    inf = g.info()
    assert [inf["tri_index_size"], inf["sorted_tris_size"]] == [o.info["tindex_size"], o.info["sorted_size"]], what


# ---- inverse-render terrain ---------------------------------------------------------

@pytest.fixture(scope="module")
def terrain():
    import torch
    print("[terrain] generating 2048^2 heightfield", flush=True)
    v, t, uv = S.terrain(2048, seed=4)
    g = rtx.build_device(torch.from_numpy(v).cuda(), torch.from_numpy(t.view(np.int32)).cuda())
    torch.cuda.synchronize()
    print(f"[terrain] {len(t)} triangles built on the GPU; oracle build", flush=True)
    o = O.build(v, t)
    return v, t, uv, g, o


def test_terrain_tree_bit_exact(terrain):
    v, t, _, g, o = terrain
    assert len(t) == 8_380_418
    codes, srt = g.morton()
    ocodes, obox = O.morton(v, t)
    assert np.array_equal(codes, ocodes)
    assert np.array_equal(srt, o.sorted)
    assert_same_nodes(g.export_bfs(), o.bfs_nodes, "terrain builder order")
    _same_tree(g, o, "terrain")
    print(f"[terrain] {o.info['nodes']} nodes, {o.info['spatial']} spatial splits, depth {o.info['depth']}", flush=True)


def test_terrain_render_matches_oracle(terrain):
    import torch
    v, t, uv, g, o = terrain
    cam, sky, sun = S.terrain_setup()
    tex = S.albedo_texture(512, seed=4)
    W, H, spp, mb = 96, 54, 2, 8
    p = rtx.make_params(W, H, spp, mb, 99, seed=4, env_mode=3, env_color=sky, uv=_tensor(uv), tex=_tensor(tex), sun=sun)
    img = rtx.render(g, cam, p).cpu().numpy()
    op = O.make_params(W, H, spp, mb, 99, seed=4, env_mode=3, env_color=sky, uv=uv, tex=tex, sun=sun)
    ref = O.render(o, cam, op, nthreads=16)
    assert img.max() > 0.05
    _close(img, ref, "terrain 96x54 2 spp depth 8")
    del torch


def test_terrain_traces_wave_per_ray_and_persistent(terrain):
    """The terrain's 28-level tree: camera rays and rays scattered over the terrain box give
    the same hits in one large call (sorted persistent traversal) and in calls of 4,096 rays
    (one wave per ray), and a subset equals the oracle's."""
    import torch
    v, t, uv, g, o = terrain
    cam, _, _ = S.terrain_setup()
    cr = S.pinhole_rays(cam, 160, 128)
    sc = S.incoherent_rays(20480, seed=12)
    lo, hi = v[:, :3].min(0), v[:, :3].max(0)
    sc[:, :3] = lo + (hi - lo) * sc[:, :3]
    rays = np.concatenate([cr, sc])
    rd = torch.from_numpy(rays).cuda()
    big = torch.zeros((len(rays), 4), device="cuda")
    small = torch.zeros((len(rays), 4), device="cuda")
    g.trace_closest_device(rd, big)
    for a in range(0, len(rays), 4096):
        g.trace_closest_device(rd[a:a + 4096], small[a:a + 4096])
    torch.cuda.synchronize()
    hb, hs = big.cpu().numpy(), small.cpu().numpy()
    assert np.array_equal(hb.view(np.uint32), hs.view(np.uint32))
    sel = np.r_[0:1000, len(cr):len(cr) + 1000]
    ref = O.trace(o, rays[sel], nthreads=16)
    assert np.array_equal(hs[sel, 3].view(np.uint32), ref["tri"])
    assert np.array_equal(hs[sel, 0].view(np.uint32), ref["t"].view(np.uint32))
    assert (ref["tri"] != 0xFFFFFFFF).mean() > 0.5


# ---- batched scenes -------------------------------------------------------------------

@pytest.fixture(scope="module")
def batched():
    print("[batched] generating 64 scenes", flush=True)
    sc = S.batched_scenes(64, seed=2)
    bvhs = [rtx.build(s["verts"], s["tris"]) for s in sc]
    return sc, bvhs


def test_batched_scene_trees_bit_exact(batched):
    sc, bvhs = batched
    with ThreadPoolExecutor(16) as ex:
        trees = list(ex.map(lambda s: O.build(s["verts"], s["tris"]), sc))
    for i, (g, o) in enumerate(zip(bvhs, trees)):
        _same_tree(g, o, f"batched scene {i} ({len(sc[i]['tris'])} triangles)")
    print(f"[batched] 64 trees bit-exact, {sum(len(s['tris']) for s in sc)} triangles", flush=True)


def test_batched_scene_renders_match_oracle(batched):
    sc, bvhs = batched
    order = np.argsort([len(s["tris"]) for s in sc])
    for i in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        s = sc[i]
        p = rtx.make_params(48, 48, 2, s["max_bounce"], 3, seed=2 + i)
        op = O.make_params(48, 48, 2, s["max_bounce"], 3, seed=2 + i)
        img = rtx.render(bvhs[i], s["cam"], p).cpu().numpy()
        _close(img, O.render(O.build(s["verts"], s["tris"]), s["cam"], op, nthreads=16),
               f"batched scene {i} ({len(s['tris'])} triangles)")


def test_batched_render_full_size_equals_plain(batched):
    """The real batched workload (64 scenes at their resolutions and spp): the hipGraph
    launch (mode 1) and the shared launches (mode 2) give the plain render (mode 0) bit
    for bit."""
    import torch
    sc, bvhs = batched
    params = [rtx.make_params(s["width"], s["height"], s["spp"], s["max_bounce"], 3, seed=2 + i)
              for i, s in enumerate(sc)]
    out = {}
    for mode in (0, 1, 2):
        imgs = [torch.full((s["height"], s["width"], 3), -1.0, device="cuda") for s in sc]
        b = rtx.Batch(bvhs, [s["cam"] for s in sc], params, [None] * 64, [None] * 64, imgs, mode)
        b.launch()
        torch.cuda.synchronize()
        out[mode] = imgs
        del b
        print(f"[batched] mode {mode} rendered", flush=True)
    for mode in (1, 2):
        for k, (a, r) in enumerate(zip(out[mode], out[0])):
            assert torch.equal(a, r), f"mode {mode} scene {k} differs from the plain render"
    assert all(float(im.min()) >= 0.0 for im in out[0])


# ---- differentiable render -----------------------------------------------------------

def test_differentiable_gradients_64_uniform_parameters():
    """Adjoint gradients of sum((I - T)^2) at 512 x 512 (4 spp, depth 4, no roulette) for
    64 albedo parameters drawn uniformly over all 4118 x 3, and the 6 emission parameters of
    the emitter, against central differences (eps 1e-3) of the oracle's float64 loss.
    A parameter passes when |adjoint - FD| <= 1e-2 |FD| + 1e-4 max|gradient| (the floor
    covers parameters whose gradient cancels to nearly zero in float32 sums)."""
    v, t, albedo_true, emission, cam = S.diff_scene(seed=3)
    W, H, spp, mb = 512, 512, 4, 4
    params = rtx.make_params(W, H, spp, mb, 99, seed=3)
    oparams = O.make_params(W, H, spp, mb, 99, seed=3)
    g = rtx.build(v, t)
    o = O.build(v, t)
    target = rtx.render(g, cam, params, _tensor(albedo_true), _tensor(emission))
    rng = np.random.default_rng(7)
    albedo = np.clip(albedo_true + rng.normal(0, 0.1, albedo_true.shape), 0.05, 0.95).astype(np.float32)
    a_t = _tensor(albedo).requires_grad_(True)
    e_t = _tensor(emission).requires_grad_(True)
    loss = ((rtx.differentiable_render(a_t, e_t, g, cam, params) - target) ** 2).sum()
    loss.backward()
    ga, ge = a_t.grad.cpu().numpy(), e_t.grad.cpu().numpy()
    tgt = target.cpu().numpy()
    floor = 1e-4 * max(np.abs(ga).max(), np.abs(ge).max())
    eps = 1e-3

    def fd(which, ti, c):
        ap, am = albedo.copy(), albedo.copy()
        ep, em = emission.copy(), emission.copy()
        (ap if which == "a" else ep)[ti, c] += eps
        (am if which == "a" else em)[ti, c] -= eps
        return (O.render_loss_f64(o, cam, oparams, ap, ep, tgt, nthreads=16) -
                O.render_loss_f64(o, cam, oparams, am, em, tgt, nthreads=16)) / (2 * eps)

    picks = [("a",) + divmod(int(f), 3) for f in rng.choice(albedo.size, 64, replace=False)]
    picks += [("e", ti, c) for ti in (20, 21) for c in range(3)]
    strict, worst = 0, 0.0
    for which, ti, c in picks:
        gv = float((ga if which == "a" else ge)[ti, c])
        f = fd(which, ti, c)
        err = abs(gv - f)
        assert err <= 1e-2 * abs(f) + floor, f"{which}[{ti},{c}]: adjoint {gv:.6g} vs FD {f:.6g}"
        if err <= 1e-2 * abs(f):
            strict += 1
        if abs(f) > floor:
            worst = max(worst, err / abs(f))
    print(f"[adjoint 512x512] {strict}/{len(picks)} parameters within 1e-2 relative; worst relative error "
          f"{worst:.2e} over parameters with |FD| above the floor {floor:.2e}", flush=True)
