"""Terrain inverse render on the GPU: textured terrain under sun + sky equals the oracle
image; the path-replay gradient per texel matches central finite differences of the
oracle's float64 loss (LDS and global accumulation paths); Adam on the texture through
the autograd function lowers the loss."""
import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _setup(n=33, tex_res=16, seed=4):
    v, t, uv = S.terrain(n=n, seed=seed)
    cam, sky, sun = S.terrain_setup()
    tex = S.albedo_texture(tex_res, seed=seed)
    return v, t, uv, tex, cam, sky, sun


@pytest.mark.parametrize("rr", [99, 2])
def test_textured_terrain_matches_oracle(rr):
    v, t, uv, tex, cam, sky, sun = _setup()
    g, o = rtx.build(v, t), O.build(v, t)
    p = rtx.make_params(48, 32, 4, 3, rr, seed=7, env_mode=3, env_color=sky, uv=_t(uv), tex=_t(tex), sun=sun)
    op = O.make_params(48, 32, 4, 3, rr, seed=7, env_mode=3, env_color=sky, uv=uv, tex=tex, sun=sun)
    img = rtx.render(g, cam, p).cpu().numpy()
    ref = O.render(o, cam, op, nthreads=16)
    bit = np.mean(img.view(np.uint32) == ref.view(np.uint32))
    print(f"[texture parity] rr={rr}: {bit * 100:.3f}% bit-identical", flush=True)
    np.testing.assert_allclose(img, ref, rtol=RTOL, atol=1e-6)
    assert img.max() > 1.0   # the sun is seen through bounces


# 64 and 6400 texels: gradients accumulated in LDS (up to 160 KiB per block); 65536: global atomics
@pytest.mark.parametrize("tex_res", [8, 80, 256])
def test_texture_gradient_matches_finite_differences(tex_res):
    import torch
    v, t, uv, tex_true, cam, sky, sun = _setup(tex_res=tex_res)
    g, o = rtx.build(v, t), O.build(v, t)
    W, H, spp, mb = 32, 24, 4, 3
    uv_t = _t(uv)
    p = rtx.make_params(W, H, spp, mb, 99, seed=3, env_mode=3, env_color=sky, uv=uv_t, tex=_t(tex_true), sun=sun)
    target = rtx.render(g, cam, p)
    rng = np.random.default_rng(1)
    tex = np.clip(tex_true + rng.normal(0, 0.15, tex_true.shape), 0.05, 0.95).astype(np.float32)
    x = _t(tex).requires_grad_(True)
    img = rtx.differentiable_render_texture(x, uv_t, g, cam, p)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    gt = x.grad.cpu().numpy()
    tgt = target.cpu().numpy()
    flat = np.argsort(-np.abs(gt).ravel())[:6]
    eps = 1e-3
    worst = 0.0
    for f in flat:
        iy, ix, c = np.unravel_index(int(f), gt.shape)
        tp, tm = tex.copy(), tex.copy()
        tp[iy, ix, c] += eps
        tm[iy, ix, c] -= eps
        lp = O.render_loss_f64(o, cam, O.make_params(W, H, spp, mb, 99, seed=3, env_mode=3, env_color=sky, uv=uv,
                                                     tex=tp, sun=sun), None, None, tgt, nthreads=16)
        lm = O.render_loss_f64(o, cam, O.make_params(W, H, spp, mb, 99, seed=3, env_mode=3, env_color=sky, uv=uv,
                                                     tex=tm, sun=sun), None, None, tgt, nthreads=16)
        fd = (lp - lm) / (2 * eps)
        err = abs(gt[iy, ix, c] - fd) / max(abs(fd), 1e-12)
        worst = max(worst, err)
        assert err < 1e-2, f"texel ({ix},{iy},{c}): adjoint {gt[iy, ix, c]:.6g} vs FD {fd:.6g}"
    print(f"[texture adjoint] {tex_res}^2 texels: worst relative error {worst:.2e}", flush=True)


def test_zero_texel_gradient_is_finite():
    """A texel channel at exactly 0 (where a clamped optimiser can put it) gets gradient 0
    from the vertices on it instead of 0/0, so the whole gradient stays finite; the other
    channels still follow the path-replay formula."""
    import torch
    v, t, uv, tex_true, cam, sky, sun = _setup(tex_res=8)
    g = rtx.build(v, t)
    uv_t = _t(uv)
    p = rtx.make_params(32, 24, 4, 3, 99, seed=3, env_mode=3, env_color=sky, uv=uv_t, tex=_t(tex_true), sun=sun)
    target = rtx.render(g, cam, p)
    tex = np.full_like(tex_true, 0.5)
    tex[::2, :, 0] = 0.0
    tex[:, ::3, 2] = 0.0
    x = _t(tex).requires_grad_(True)
    loss = ((rtx.differentiable_render_texture(x, uv_t, g, cam, p) - target) ** 2).sum()
    loss.backward()
    gr = x.grad.cpu().numpy()
    assert np.isfinite(gr).all()
    assert np.all(gr[tex == 0.0] == 0.0)
    assert np.abs(gr[tex != 0.0]).max() > 0.0


def test_adam_on_texture_lowers_loss():
    import torch
    v, t, uv, tex_true, cam, sky, sun = _setup(tex_res=16)
    g = rtx.build(v, t)
    uv_t = _t(uv)
    p = rtx.make_params(48, 32, 8, 3, 99, seed=5, env_mode=3, env_color=sky, uv=uv_t, tex=_t(tex_true), sun=sun)
    target = rtx.render(g, cam, p).clone()
    x = torch.full_like(_t(tex_true), 0.5).requires_grad_(True)
    opt = torch.optim.Adam([x], lr=2e-2)
    losses = []
    for _ in range(25):
        opt.zero_grad()
        loss = ((rtx.differentiable_render_texture(x, uv_t, g, cam, p) - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            x.clamp_(1e-3, 1.0)
        losses.append(float(loss.detach()))
    print(f"[texture adam] loss {losses[0]:.4g} -> {losses[-1]:.4g}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
