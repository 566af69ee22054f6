"""Terrain inverse-render extension on the CPU side: the heightfield generator, and
oracle identities of texture albedo and the sun + sky environment (env_mode 3)."""
import numpy as np

from oracle import oracle as O
from rtx import scenes as S


def test_terrain_generator():
    """2 (n-1)^2 triangles over an n x n grid, wound so the reference's back-face test
    keeps them for rays from above, uv = (x, z) in [0,1], unit normals pointing up; the
    camera of the inverse render sees mostly terrain."""
    v, t, uv = S.terrain(n=33, seed=4)
    assert v.shape == (33 * 33, 6) and t.shape == (2 * 32 * 32, 3) and uv.shape == (33 * 33, 2)
    p = v[:, :3][t]
    cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert (cr[:, 1] < 0).all()
    assert uv.min() == 0.0 and uv.max() == 1.0
    np.testing.assert_allclose(np.linalg.norm(v[:, 3:], axis=1), 1.0, atol=1e-6)
    assert (v[:, 4] > 0).all()
    cam, _, _ = S.terrain_setup()
    h = O.trace(O.build(v, t), S.pinhole_rays(cam, 64, 36))
    assert (h["tri"] != 0xFFFFFFFF).mean() > 0.8


def test_constant_texture_equals_constant_albedo():
    """A texture filled with one colour renders exactly like that colour as per-triangle
    albedo, and a sun of zero radiance equals the constant environment."""
    v, t, uv = S.terrain(n=17, seed=5)
    o = O.build(v, t)
    cam, sky, sun = S.terrain_setup()
    col = np.array([0.3, 0.6, 0.45], np.float32)
    tex = np.tile(col, (8, 8, 1))
    alb = np.tile(col, (len(t), 1))
    a = O.render(o, cam, O.make_params(24, 16, 2, 3, 99, seed=2, env_mode=3, env_color=sky, uv=uv, tex=tex,
                                       sun=sun), nthreads=8)
    b = O.render(o, cam, O.make_params(24, 16, 2, 3, 99, seed=2, env_mode=3, env_color=sky, sun=sun), alb,
                 nthreads=8)
    assert np.array_equal(a, b)
    dark = (sun[0], sun[1], (0.0, 0.0, 0.0))
    c = O.render(o, cam, O.make_params(24, 16, 2, 3, 99, seed=2, env_mode=3, env_color=sky, sun=dark), alb,
                 nthreads=8)
    d = O.render(o, cam, O.make_params(24, 16, 2, 3, 99, seed=2, env_mode=1, env_color=sky), alb, nthreads=8)
    assert np.array_equal(c, d)
    assert not np.array_equal(b, d)   # the sun is seen


def test_texture_lookup_reads_the_hit_texel():
    """A texture that is black except one texel: only pixels whose primary hit lands in that
    texel's uv square (and their bounce light) change."""
    v, t, uv = S.terrain(n=9, seed=6, height=0.0)   # flat: uv is the hit's (x, z)
    o = O.build(v, t)
    cam = S.camera_lookat((0.5, 1.0, 0.5001), (0.5, 0.0, 0.5), 40.0)
    tex = np.zeros((4, 4, 3), np.float32)
    base = O.render(o, cam, O.make_params(32, 32, 1, 1, 99, seed=1, env_mode=1, env_color=(1, 1, 1), uv=uv,
                                          tex=tex), nthreads=4)
    assert (base == 0).all()                     # black ground, the bounce is not taken
    tex[2, 1] = 1.0                              # texel x = 1, y = 2: u in [0.25, 0.5), v in [0.5, 0.75)
    img = O.render(o, cam, O.make_params(32, 32, 1, 1, 99, seed=1, env_mode=1, env_color=(1, 1, 1), uv=uv,
                                         tex=tex), nthreads=4)
    rays = S.pinhole_rays(cam, 32, 32)
    hit = O.trace(o, rays)
    p = rays[:, :3] + hit["t"][:, None] * rays[:, 4:7]
    inside = (p[:, 0] >= 0.26) & (p[:, 0] < 0.49) & (p[:, 2] >= 0.51) & (p[:, 2] < 0.74)
    outside = (p[:, 0] < 0.24) | (p[:, 0] > 0.51) | (p[:, 2] < 0.49) | (p[:, 2] > 0.76)
    lit = img.reshape(-1, 3).sum(1) > 0
    assert inside.any() and lit[inside].all() and not lit[outside].any()
