"""GPU path tracer parity: images equal the oracle's restatement of the reference loop
(same sample table) within 1e-4 relative per pixel; batched and graph-replayed renders
equal single-scene renders; adjoint gradients match central finite differences."""
import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S
from scenes_util import reference_test_scene

pytestmark = pytest.mark.gpu

RTOL = 1e-4   # per-pixel relative tolerance of the image parity contract


def _tensor(a):
    import torch
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _close(gpu, orc, what):
    bit = np.mean(gpu.view(np.uint32) == orc.view(np.uint32))
    print(f"[render parity] {what}: {bit * 100:.3f}% of pixel channels bit-identical", flush=True)
    np.testing.assert_allclose(gpu, orc, rtol=RTOL, atol=1e-6, err_msg=what)


def _render_both(v, t, albedo, emission, cam, params, oparams):
    g = rtx.build(v, t)
    o = O.build(v, t)
    img = rtx.render(g, cam, params, _tensor(albedo), _tensor(emission)).cpu().numpy()
    ref = O.render(o, cam, oparams, albedo, emission, nthreads=16)
    return img, ref


@pytest.mark.parametrize("spp,max_bounce,rr", [(4, 5, 3), (2, 8, 3), (3, 2, 99)])
def test_cornell_matches_oracle(spp, max_bounce, rr):
    v, t, albedo, emission, cam = S.cornell_box()
    p = rtx.make_params(64, 48, spp, max_bounce, rr, seed=0)
    op = O.make_params(64, 48, spp, max_bounce, rr, seed=0)
    img, ref = _render_both(v, t, albedo, emission, cam, p, op)
    assert img.shape == (48, 64, 3) and np.isfinite(img).all()
    assert img.max() > 0.05
    _close(img, ref, f"cornell spp={spp} maxBounce={max_bounce} rr={rr}")


def test_reference_shading_defaults():
    """No materials: the reference's constant 0.8 albedo and sky gradient (main.cpp:160-179)."""
    v, t = reference_test_scene(rtx.geo_sphere, rtx.geo_box)
    cam = S.camera_lookat((0, 2, -12), (0, 0, 4), 55.0)
    p = rtx.make_params(80, 60, 3, 8, 3, seed=5)
    op = O.make_params(80, 60, 3, 8, 3, seed=5)
    img, ref = _render_both(v, t, None, None, cam, p, op)
    _close(img, ref, "reference defaults")


def test_constant_environment():
    v, t, albedo, _, cam = S.cornell_box(seed=4)
    p = rtx.make_params(40, 40, 2, 4, 3, seed=9, env_mode=1, env_color=(0.3, 0.5, 0.9))
    op = O.make_params(40, 40, 2, 4, 3, seed=9, env_mode=1, env_color=(0.3, 0.5, 0.9))
    img, ref = _render_both(v, t, albedo, None, cam, p, op)
    _close(img, ref, "constant environment")


def test_sky_closed_form():
    """One triangle behind the camera: every pixel is the mean over samples of the sky
    gradient (1,1,1)(1-s) + (.5,.7,1)s with s = dir.y/2 + 1/2 of its primary ray."""
    v = np.array([[0, 0, 10, 0, 0, -1], [1, 0, 10, 0, 0, -1], [0, 1, 10, 0, 0, -1]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    cam = S.camera_lookat((0, 0, 0), (0, 0, -1), 60.0)
    W, H, spp = 24, 16, 5
    g = rtx.build(v, t)
    img = rtx.render(g, cam, rtx.make_params(W, H, spp, 3, 3, seed=2)).cpu().numpy().astype(np.float64)
    exp = np.zeros((H, W, 3))
    for y in range(H):
        for x in range(W):
            for s in range(spp):
                pix = y * W + x
                jx, jy = O.sample(2, pix, s, 0), O.sample(2, pix, s, 1)
                u = ((x + jx) / W * 2 - 1) * (W / H)
                vv = (y + jy) / H * 2 - 1
                d = cam["fwd"] + cam["right"] * u * cam["tan_half_fov_y"] + cam["up"] * vv * cam["tan_half_fov_y"]
                d = d / np.linalg.norm(d)
                sy = d[1] * 0.5 + 0.5
                exp[y, x] += np.array([1 - sy + 0.5 * sy, 1 - sy + 0.7 * sy, 1.0]) / spp
    np.testing.assert_allclose(img, exp, rtol=2e-6, atol=2e-6)


def _small_batch_scenes():
    out = []
    for i, (rings, res, spp) in enumerate([(8, 24, 3), (20, 32, 2), (12, 16, 5)]):
        v, t = S.uv_sphere(rings, 2 * rings, 1.0, (0, 0, 0), 0.05, seed=i)
        cam = S.camera_lookat((0.3 * i, 0.4, -3.2), (0, 0, 0), 45.0)
        out.append((v, t, cam, rtx.make_params(res, res, spp, 6, 3, seed=i), O.make_params(res, res, spp, 6, 3, seed=i)))
    return out


def test_single_scene_render_variants_bit_identical():
    """One scene per launch sequence: the traversal keeps the tree in scalar registers. The
    persistent kernel (render_loop 2 on a batch of at least trace_grid_blocks * 512 paths),
    the chunked loops and the single-loop form give bit-identical images."""
    v, t, cam, p, _ = _small_batch_scenes()[1]   # 32 x 32 x 2 spp = 2048 paths
    g = rtx.build(v, t)
    saved = rtx.get_tuning()
    try:
        ref = None
        for loop, coop, pgrid, bp in ((1, 1, 2048, 1 << 22), (2, 1, 2, 1 << 22), (2, 0, 1, 1024), (0, 0, 2048, 1 << 22),
                                      (2, 1, 2048, 1 << 22)):
            rtx.set_tuning(render_loop=loop, render_leaf_coop=coop, trace_grid_blocks=pgrid, batch_paths=bp)
            print(f"[single-scene variant] loop={loop} coop={coop} pgrid={pgrid} batch={bp}", flush=True)
            img = rtx.render(g, cam, p).cpu().numpy()
            if ref is None:
                ref = img
            assert np.array_equal(img, ref), (loop, coop, pgrid, bp)
    finally:
        rtx.set_tuning(**saved)


def test_batched_modes_match_single_and_oracle():
    import torch
    sc = _small_batch_scenes()
    bvhs = [rtx.build(v, t) for v, t, _, _, _ in sc]
    singles = [rtx.render(b, c, p).cpu().numpy() for b, (_, _, c, p, _) in zip(bvhs, sc)]
    for (v, t, c, _, op), img in zip(sc, singles):
        _close(img, O.render(O.build(v, t), c, op, nthreads=16), "batch scene vs oracle")
    for mode in (0, 1, 2):
        imgs = [torch.zeros((p.height, p.width, 3), device="cuda") for _, _, _, p, _ in sc]
        b = rtx.Batch(bvhs, [c for _, _, c, _, _ in sc], [p for _, _, _, p, _ in sc], [None] * 3, [None] * 3,
                      imgs, mode)
        b.launch()
        torch.cuda.synchronize()
        for a, s in zip(imgs, singles):
            assert np.array_equal(a.cpu().numpy(), s), f"mode {mode} differs from single-scene render"


def test_batch_of_more_scenes_than_the_lds_table():
    """A shared batch of 70 scenes (more than the 64 the renderer's LDS scene table holds,
    so lanes read their trees through global scene descriptors) and of 20 scenes (the LDS
    table) give the single-scene images, with the chunked and the persistent kernel."""
    import torch
    sc = []
    for i in range(70):
        v, t = S.uv_sphere(4 + i % 5, 8 + 2 * (i % 5), 1.0, (0, 0, 0), 0.05, seed=100 + i)
        cam = S.camera_lookat((0.05 * (i % 7), 0.3, -3.0), (0, 0, 0), 45.0)
        sc.append((v, t, cam, rtx.make_params(8, 8, 2, 4, 3, seed=i)))
    bvhs = [rtx.build(v, t) for v, t, _, _ in sc]
    singles = [rtx.render(b, c, p).cpu().numpy() for b, (_, _, c, p) in zip(bvhs, sc)]
    saved = rtx.get_tuning()
    try:
        for loop, pgrid in ((1, 2048), (2, 1)):
            rtx.set_tuning(render_loop=loop, trace_grid_blocks=pgrid)
            for count in (70, 20):
                print(f"[many scenes] loop={loop} scenes={count}", flush=True)
                part = sc[:count]
                imgs = [torch.zeros((p.height, p.width, 3), device="cuda") for _, _, _, p in part]
                b = rtx.Batch(bvhs[:count], [c for _, _, c, _ in part], [p for _, _, _, p in part], [None] * count,
                              [None] * count, imgs, 2)
                b.launch()
                torch.cuda.synchronize()
                for k, (a, ref) in enumerate(zip(imgs, singles)):
                    assert np.array_equal(a.cpu().numpy(), ref), (loop, count, k)
                del b
    finally:
        rtx.set_tuning(**saved)


def test_render_trace_variants_bit_identical():
    """The wavefront renderer's traversal variants (one loop, while-while with any
    render_wait_lanes, cooperative leaves, the persistent kernel, stack depths, grid and
    batch sizes) give bit-identical images in a shared multi-scene batch."""
    import torch
    sc = _small_batch_scenes()
    bvhs = [rtx.build(v, t) for v, t, _, _, _ in sc]
    saved = rtx.get_tuning()
    try:
        ref = None
        for loop, depth, wait, grid, bp, coop, pgrid in (
                (1, 8, 0, 2048, 1 << 22, 0, 2048), (1, 12, 16, 2048, 1 << 22, 0, 2048),
                (1, 8, 8, 8, 4096, 0, 2048), (1, 16, 32, 64, 8192, 0, 2048),
                (0, 8, 0, 512, 1 << 20, 0, 2048), (1, 8, 0, 2048, 1 << 22, 1, 2048),
                (1, 12, 32, 8, 4096, 1, 2048), (0, 8, 0, 512, 1 << 20, 1, 2048),
                (2, 8, 8, 16, 1 << 22, 1, 2), (2, 12, 0, 16, 4096, 0, 2), (2, 16, 4, 16, 2048, 1, 1),
                (2, 8, 16, 2048, 1 << 20, 1, 2048)):
            # render_loop 2 runs the persistent kernel on batches of at least
            # trace_grid_blocks * 4 waves * 128 paths (2 blocks: 1024 paths, every lane takes
            # many paths) and the chunked kernel on smaller ones (the last variant)
            rtx.set_tuning(render_loop=loop, render_lds_depth=depth, render_wait_lanes=wait, grid_blocks=grid,
                           batch_paths=bp, render_leaf_coop=coop, trace_grid_blocks=pgrid)
            print(f"[render variant] loop={loop} depth={depth} wait={wait} grid={grid} coop={coop} pgrid={pgrid}",
                  flush=True)
            imgs = [torch.zeros((p.height, p.width, 3), device="cuda") for _, _, _, p, _ in sc]
            b = rtx.Batch(bvhs, [c for _, _, c, _, _ in sc], [p for _, _, _, p, _ in sc], [None] * 3, [None] * 3,
                          imgs, 2)
            b.launch()
            torch.cuda.synchronize()
            out = [a.cpu().numpy() for a in imgs]
            if ref is None:
                ref = out
            for a, r in zip(out, ref):
                assert np.array_equal(a, r), (loop, depth, wait, grid, coop)
    finally:
        rtx.set_tuning(**saved)


def test_graph_replays_match_eager():
    """The captured batch is replayed several times with eager work in between; every
    replay must reproduce the eager render."""
    import torch
    sc = _small_batch_scenes()
    bvhs = [rtx.build(v, t) for v, t, _, _, _ in sc]
    eager = [rtx.render(b, c, p).cpu().numpy() for b, (_, _, c, p, _) in zip(bvhs, sc)]
    imgs = [torch.zeros((p.height, p.width, 3), device="cuda") for _, _, _, p, _ in sc]
    b = rtx.Batch(bvhs, [c for _, _, c, _, _ in sc], [p for _, _, _, p, _ in sc], [None] * 3, [None] * 3, imgs, 1)
    for rep in range(3):
        for im in imgs:
            im.fill_(-1.0)
        b.launch()
        torch.cuda.synchronize()
        for a, e in zip(imgs, eager):
            assert np.array_equal(a.cpu().numpy(), e), f"replay {rep} differs"
        _ = rtx.render(bvhs[0], sc[0][2], sc[0][3])   # eager work between replays
        torch.cuda.synchronize()


def test_render_is_stream_ordered_and_refuses_capture():
    """rtx.render returns before its work is done (stream-ordered); renders queued back to
    back on one stream, with their path state recycled through the scratch cache, equal the
    same renders run one at a time; and a render into a capturing stream is refused (its path
    state goes back to the cache once queued, so a replay would use blocks of later calls)."""
    import torch
    sc = _small_batch_scenes()
    bvhs = [rtx.build(v, t) for v, t, _, _, _ in sc]
    one = []
    for b, (_, _, c, p, _) in zip(bvhs, sc):
        one.append(rtx.render(b, c, p).cpu().numpy())
        torch.cuda.synchronize()
    s = torch.cuda.Stream()
    outs = []
    with torch.cuda.stream(s):
        for rep in range(4):
            for b, (_, _, c, p, _) in zip(bvhs, sc):
                outs.append(rtx.render(b, c, p, stream=s))
    s.synchronize()
    for k, o in enumerate(outs):
        assert np.array_equal(o.cpu().numpy(), one[k % 3]), f"queued render {k} differs"
    v, t, c, p, _ = sc[0]
    img = torch.zeros((p.height, p.width, 3), device="cuda")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        g.capture_begin()
        try:
            with pytest.raises(rtx.RtxError, match="capture"):
                rtx.render(bvhs[0], c, p, out=img, stream=s)
            img.add_(1.0)   # some captured work
        finally:
            g.capture_end()
    s.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert float(img.min()) == 1.0 and float(img.max()) == 1.0


def test_adjoint_matches_finite_differences():
    """Path-replay gradients of sum((I - T)^2) w.r.t. per-triangle albedo and emission
    against central differences (eps 1e-3) of the oracle renderer, 1e-2 relative."""
    import torch
    v, t, albedo_true, emission, cam = S.diff_scene(seed=3)
    W, H, spp, mb = 40, 40, 4, 4
    params = rtx.make_params(W, H, spp, mb, 99, seed=3)
    oparams = O.make_params(W, H, spp, mb, 99, seed=3)
    g = rtx.build(v, t)
    o = O.build(v, t)
    target = rtx.render(g, cam, params, _tensor(albedo_true), _tensor(emission))
    rng = np.random.default_rng(0)
    albedo = np.clip(albedo_true + rng.normal(0, 0.1, albedo_true.shape), 0.05, 0.95).astype(np.float32)
    a_t = _tensor(albedo).requires_grad_(True)
    e_t = _tensor(emission).requires_grad_(True)
    img = rtx.differentiable_render(a_t, e_t, g, cam, params)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    ga = a_t.grad.cpu().numpy()
    ge = e_t.grad.cpu().numpy()
    tgt = target.cpu().numpy()
    mag = np.abs(ga).ravel()
    cand = np.argsort(-mag)[:400]
    picks = list(rng.choice(cand, 14, replace=False))
    eps = 1e-3
    worst = 0.0
    for flat in picks:
        ti, c = divmod(int(flat), 3)
        ap, am = albedo.copy(), albedo.copy()
        ap[ti, c] += eps
        am[ti, c] -= eps
        fd = (O.render_loss_f64(o, cam, oparams, ap, emission, tgt, nthreads=16) -
              O.render_loss_f64(o, cam, oparams, am, emission, tgt, nthreads=16)) / (2 * eps)
        err = abs(ga[ti, c] - fd) / max(abs(fd), 1e-12)
        worst = max(worst, err)
        assert err < 1e-2, f"albedo[{ti},{c}]: adjoint {ga[ti, c]:.6g} vs FD {fd:.6g}"
    for ti, c in [(20, 0), (21, 1)]:
        ep, em = emission.copy(), emission.copy()
        ep[ti, c] += eps
        em[ti, c] -= eps
        fd = (O.render_loss_f64(o, cam, oparams, albedo, ep, tgt, nthreads=16) -
              O.render_loss_f64(o, cam, oparams, albedo, em, tgt, nthreads=16)) / (2 * eps)
        assert abs(ge[ti, c] - fd) <= 1e-2 * abs(fd), f"emission[{ti},{c}]: {ge[ti, c]} vs FD {fd}"
    print(f"[adjoint] worst relative error over {len(picks)} albedo parameters: {worst:.2e}", flush=True)


def test_adjoint_rejects_roulette():
    import torch
    v, t, albedo, emission, cam = S.diff_scene(seed=3)
    g = rtx.build(v, t)
    p = rtx.make_params(8, 8, 1, 4, 3, seed=1)
    with pytest.raises(rtx.RtxError, match="rr_bounce"):
        rtx.render_adjoint(g, cam, p, _tensor(albedo), _tensor(emission), torch.zeros((8, 8, 3), device="cuda"))


def test_ibl_environment_matches_oracle():
    """Equirectangular IBL miss shading (IBL::sample, main.cpp:107-117) on both sides."""
    import torch
    v, t, albedo, _, cam = S.cornell_box(seed=6)
    rng = np.random.default_rng(3)
    ibl = (rng.random((32, 64, 3)) * 2).astype(np.float32)
    p = rtx.make_params(48, 40, 2, 4, 3, seed=4, env_mode=2, ibl=torch.from_numpy(ibl).cuda())
    op = O.make_params(48, 40, 2, 4, 3, seed=4, env_mode=2, ibl=ibl)
    img, ref = _render_both(v, t, albedo, None, cam, p, op)
    _close(img, ref, "IBL environment")


def test_random_walk_closed_forms_on_gpu():
    """The GPU renderer's bounce sampling (cosine lobe, toWorld frame) and Russian roulette
    (survival max(throughput), throughput /= p) against the closed forms of
    tests/closed_forms.py; every image also equals the oracle's within the image tolerance."""
    import torch
    import closed_forms as CF

    def both(v, t, albedo, emission, cam, W, H, spp, mb, rr, seed, env_mode=0, ibl=None):
        g = rtx.build(v, t)
        ib = None if ibl is None else torch.from_numpy(ibl).cuda()
        p = rtx.make_params(W, H, spp, mb, rr, seed=seed, env_mode=env_mode, ibl=ib)
        img = rtx.render(g, cam, p, _tensor(albedo), _tensor(emission)).cpu().numpy()
        op = O.make_params(W, H, spp, mb, rr, seed=seed, env_mode=env_mode, ibl=ibl)
        _close(img, O.render(O.build(v, t), cam, op, albedo, emission, nthreads=16), f"closed-form scene mb={mb}")
        return img

    results = CF.all_checks(both)
    for r in results:
        print("[closed form]", r, flush=True)
    bad = [r for r in results if not r.ok]
    assert not bad, bad


def _camera_rays(cam, W, H, spp, seed):
    """The renderer's camera rays (primaryRay, main.cpp:69-83, with the sample table's jitter)
    in float32 and the renderer's operation order, path sample * H * W + slot (the slot's
    pixel by rtx.tile_pixel)."""
    f = np.float32
    n = W * H
    rays = np.zeros((spp * n, 8), f)
    fwd, rt, up = (np.asarray(cam[k], f) for k in ("fwd", "right", "up"))
    th = f(cam["tan_half_fov_y"])
    aspect = f(W) / f(H)
    slot_pixel = rtx.tile_pixel(np.arange(n), W, H)
    for s in range(spp):
        for slot in range(n):
            pix = int(slot_pixel[slot])
            x, y = pix % W, pix // W
            jx, jy = f(O.sample(seed, pix, s, 0)), f(O.sample(seed, pix, s, 1))
            u = ((f(x) + jx) / f(W) * f(2) - f(1)) * aspect
            v = (f(y) + jy) / f(H) * f(2) - f(1)
            d = (fwd + (u * th) * rt) + (v * th) * up
            inv = f(1) / np.sqrt((d[0] * d[0] + d[1] * d[1]) + (d[2] * d[2] + f(0)))
            rays[s * n + slot, 0:3] = cam["pos"]
            rays[s * n + slot, 3] = f(1e-3)
            rays[s * n + slot, 4:7] = inv * d
            rays[s * n + slot, 7] = np.inf
    return rays


def test_primal_render_and_path_records():
    """rtx_render_primal gives rtx_render's image bit for bit; each pixel is the mean of its
    stored path radiances summed in sample order (k_accum's order); and each record's hit is
    the oracle's closest hit of that path's camera ray, bit for bit (the replay uses it in
    place of tracing the camera rays again)."""
    v, t, albedo, emission, cam = S.diff_scene(seed=3)
    g = rtx.build(v, t)
    W, H, spp = 32, 24, 5
    p = rtx.make_params(W, H, spp, 4, 99, seed=2)
    a_t, e_t = _tensor(albedo), _tensor(emission)
    ref = rtx.render(g, cam, p, a_t, e_t).cpu().numpy()
    img, paths = rtx.render_primal(g, cam, p, a_t, e_t)
    img, rec = img.cpu().numpy(), paths.cpu().numpy()
    assert rec.shape == (2, spp * H * W, 4)
    rad = np.empty((spp, H * W, 3), np.float32)
    rad[:, rtx.tile_pixel(np.arange(H * W), W, H)] = rec[0, :, 0:3].reshape(spp, H * W, 3)   # by pixel
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    acc = np.zeros((H * W, 3), np.float32)
    for s in range(spp):
        acc = (acc + rad[s]).astype(np.float32)
    assert np.array_equal((acc / np.float32(spp)).astype(np.float32).reshape(H, W, 3).view(np.uint32),
                          ref.view(np.uint32))
    assert not rec[0, :, 3].any()
    hits = O.trace(O.build(v, t), _camera_rays(cam, W, H, spp, seed=2))
    tri = rec[1, :, 3].view(np.uint32)
    assert np.array_equal(tri, hits["tri"])
    hit = tri != 0xFFFFFFFF
    assert hit.mean() > 0.9
    for k, c in (("t", 0), ("u", 1), ("v", 2)):
        assert np.array_equal(rec[1, hit, c].view(np.uint32), hits[k][hit].view(np.uint32)), k


def test_adjoint_replay_equals_self_contained_and_memory_is_flat():
    """The replay from stored path radiance and the self-contained adjoint (its own primal
    pass) give the same gradients (up to the order of float atomics), and the adjoint's
    device scratch does not grow with max_bounce."""
    import torch
    v, t, albedo, emission, cam = S.diff_scene(seed=3)
    g = rtx.build(v, t)
    W, H, spp = 48, 40, 4
    a_t, e_t = _tensor(albedo), _tensor(emission)
    peaks = {}
    for mb in (2, 30):
        p = rtx.make_params(W, H, spp, mb, 99, seed=5)
        img, paths = rtx.render_primal(g, cam, p, a_t, e_t)
        d = (2 * (img - 0.3)).contiguous()
        ga1, ge1 = rtx.render_adjoint(g, cam, p, a_t, e_t, d, path_records=paths)
        torch.cuda.synchronize()
        rtx.debug_scratch_peak(reset=True)
        ga2, ge2 = rtx.render_adjoint(g, cam, p, a_t, e_t, d)
        torch.cuda.synchronize()
        peaks[mb] = rtx.debug_scratch_peak(reset=True)
        for x, y in ((ga1, ga2), (ge1, ge2)):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            assert np.abs(x).max() > 0
            np.testing.assert_allclose(x, y, rtol=1e-4, atol=1e-6 * np.abs(x).max())
    print(f"[adjoint memory] scratch peak {peaks[2]} B at max_bounce 2, {peaks[30]} B at max_bounce 30", flush=True)
    assert peaks[2] > 0 and abs(peaks[30] - peaks[2]) <= 4096


def test_adjoint_over_many_batches_equals_one_batch():
    """With paths processed in batches of 1024 (the library's smallest), the primal's stored
    radiance and the replay's gradient scatter index each batch's paths by their global
    number: gradients equal the one-batch run's up to the order of float atomics, for the
    per-triangle arrays (LDS form) and for texels (global-atomics form)."""
    import torch
    v, t, albedo, emission, cam = S.diff_scene(seed=3)
    g = rtx.build(v, t)
    W, H, spp = 40, 36, 3   # 4320 paths: four full batches of 1024 and a partial one
    p = rtx.make_params(W, H, spp, 3, 99, seed=9)
    a_t, e_t = _tensor(albedo), _tensor(emission)
    saved = rtx.get_tuning()
    out = {}
    try:
        for bp in (1 << 23, 1024):
            rtx.set_tuning(batch_paths=bp)
            img, paths = rtx.render_primal(g, cam, p, a_t, e_t)
            d = (2 * (img - 0.25)).contiguous()
            ga, ge = rtx.render_adjoint(g, cam, p, a_t, e_t, d, path_records=paths)
            ga2, ge2 = rtx.render_adjoint(g, cam, p, a_t, e_t, d)
            torch.cuda.synchronize()
            out[bp] = [x.cpu().numpy() for x in (img, paths, ga, ge, ga2, ge2)]
    finally:
        rtx.set_tuning(**saved)
    one, many = out[1 << 23], out[1024]
    assert np.array_equal(one[0].view(np.uint32), many[0].view(np.uint32))   # image
    assert np.array_equal(one[1].view(np.uint32), many[1].view(np.uint32))   # path radiance
    for x, y in zip(one[2:], many[2:]):
        assert np.abs(x).max() > 0
        np.testing.assert_allclose(y, x, rtol=1e-4, atol=1e-6 * np.abs(x).max())

    # texels: a 512^2 texture on a small terrain takes the global-atomics form
    tv, tt, uv = S.terrain(33, seed=4)
    tg = rtx.build(tv, tt)
    tcam, sky, sun = S.terrain_setup()
    uv_t = torch.from_numpy(uv).cuda()
    tex = torch.full((512, 512, 3), 0.5, device="cuda")
    tp = rtx.make_params(W, H, spp, 3, 99, seed=4, env_mode=3, env_color=sky, uv=uv_t, tex=tex, sun=sun)
    grads = {}
    try:
        for bp in (1 << 23, 1024):
            rtx.set_tuning(batch_paths=bp)
            img, paths = rtx.render_primal(tg, tcam, tp)
            d = (2 * (img - 0.25)).contiguous()
            gt, _ = rtx.render_adjoint(tg, tcam, tp, None, None, d, path_records=paths)
            torch.cuda.synchronize()
            grads[bp] = gt.cpu().numpy()
    finally:
        rtx.set_tuning(**saved)
    x, y = grads[1 << 23], grads[1024]
    assert np.abs(x).max() > 0
    np.testing.assert_allclose(y, x, rtol=1e-4, atol=1e-6 * np.abs(x).max())
