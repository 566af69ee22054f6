"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, the scene front-end equals the reference generators, file errors map to
the reference's failure cases, and the oracle renderer meets closed forms."""
import os
import re

import numpy as np
import pytest

import rtx
from oracle import oracle as O
from rtx import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "rtx.h")).read()
    declared = sorted(set(re.findall(r"\b(rtx_[a-z_]+)\s*\(", hdr)))
    L = rtx.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"
    assert sorted(rtx.EXPORTS) == declared


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("args", [((0, 5, 0), 2.0, 16, 32), ((1.5, -2, 3), 0.7, 33, 64), ((0, 0, 0), 1.0, 2, 3)])
def test_sphere_generator_matches_reference(args):
    a, b = rtx.geo_sphere(*args), O.ref_geo_sphere(*args)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("args", [((0, 0, 5), (1, 1, 1), (0, 0, 0, 1)), ((0, -3, 0), (20, 0.1, 20), (0.1, 0.2, 0.3, 0.927))])
def test_box_generator_matches_reference(args):
    a, b = rtx.geo_box(*args), O.ref_geo_box(*args)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])


def test_scene_file_front_end(tmp_path):
    scn = tmp_path / "s.scn"
    scn.write_text("screen 80 60\ncamera 0 1 -5 lookat 0 0 0 60\nsphere 0 1 0 1\n"
                   "box -2 1 2 1 1 1 q 0 0 0 1\nsphere 0, 0, 0, 1000\nbox 1 2 3\n")
    v, t = rtx.load_scene(scn)
    sv, st = rtx.geo_sphere((0, 1, 0), 1.0)
    bv, bt = rtx.geo_box((-2, 1, 2), (1, 1, 1))
    assert len(t) == len(st) + len(bt)       # malformed lines are skipped, as readScene does
    n = np.linalg.norm(v[:, 3:], axis=1)
    np.testing.assert_allclose(n, 1.0, atol=1e-6)


def test_scene_file_errors(tmp_path):
    with pytest.raises(rtx.RtxError, match="failed to open"):
        rtx.load_scene(tmp_path / "missing.scn")
    empty = tmp_path / "empty.scn"
    empty.write_text("screen 10 10\n")
    with pytest.raises(rtx.RtxError, match="no vertices"):
        rtx.load_scene(empty)
    mesh = tmp_path / "mesh.scn"
    mesh.write_text("mesh nothere.ply 0 0 0 1 q 0 0 0 1\n")
    with pytest.raises(rtx.RtxError, match="failed to open mesh"):
        rtx.load_scene(mesh)
    mesh.write_text("mesh dwarf.x 0 0 0 1 q 0 0 0 1\n")
    with pytest.raises(rtx.RtxError, match="ASCII .ply only"):
        rtx.load_scene(mesh)
    (tmp_path / "bin.ply").write_text("ply\nformat binary_little_endian 1.0\nelement vertex 0\nend_header\n")
    mesh.write_text("mesh bin.ply 0 0 0 1 q 0 0 0 1\n")
    with pytest.raises(rtx.RtxError, match="only ASCII PLY"):
        rtx.load_scene(mesh)


def test_ply_mesh_lines(tmp_path):
    """A quad and a triangle in ASCII PLY, translated, scaled and rotated 90 degrees about z:
    fan triangulation, smooth normals from face normals, transform translate*scale*rotate."""
    (tmp_path / "q.ply").write_text(
        "ply\nformat ascii 1.0\nelement vertex 5\nproperty float x\nproperty float y\nproperty float z\n"
        "element face 2\nproperty list uchar int vertex_indices\nend_header\n"
        "0 0 0\n1 0 0\n1 1 0\n0 1 0\n2 2 0\n4 0 1 2 3\n3 1 4 2\n")
    s = float(np.sqrt(0.5))
    (tmp_path / "m.scn").write_text(f"mesh q.ply 1 2 3 2 q 0 0 {s} {s}\n")
    v, t = rtx.load_scene(tmp_path / "m.scn")
    assert v.shape == (5, 6) and t.tolist() == [[0, 1, 2], [0, 2, 3], [1, 4, 2]]
    # rotation by 90 degrees about z maps (x, y) -> (-y, x); then scale 2, translate (1, 2, 3)
    np.testing.assert_allclose(v[:, :3], [[1, 2, 3], [1, 4, 3], [-1, 4, 3], [-1, 2, 3], [-3, 6, 3]], atol=1e-5)
    np.testing.assert_allclose(v[:, 3:], np.tile([0, 0, 1], (5, 1)), atol=1e-6)


@pytest.mark.skipif(not os.path.exists("/root/reference/bunny.ply"), reason="reference assets not present")
def test_reference_bunny_loads(tmp_path):
    (tmp_path / "b.scn").write_text("mesh /root/reference/bunny.ply 0 0 0 10 q 0 0 0 1\n")
    v, t = rtx.load_scene(tmp_path / "b.scn")
    assert v.shape == (35947, 6) and t.shape == (69451, 3)


def _golden_mesh_files(tmp_path):
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_bunny_head.npz"))
    (tmp_path / "bunny_head.ply").write_bytes(g["ply"].tobytes())
    (tmp_path / "head.scn").write_bytes(g["scn"].tobytes())
    return g


def test_ply_golden_mesh_loads_as_recorded(tmp_path):
    """The PLY-derived mesh of reference_bunny_head.npz (the top of the reference's bunny.ply
    under a translate * scale * rotate `mesh` line) loads into the recorded arrays bit for
    bit, and agrees with an independent float64 model of the loader: positions
    t + s R(q) p, smooth normals = R(q) times the sum of the unit face normals, renormalized."""
    g = _golden_mesh_files(tmp_path)
    v, t = rtx.load_scene(tmp_path / "head.scn")
    assert np.array_equal(v.view(np.uint32), g["verts"].view(np.uint32)) and np.array_equal(t, g["tris"])
    lines = g["ply"].tobytes().decode().split("\n")
    end = lines.index("end_header")
    nv = int(lines[2].split()[2])
    P = np.array([list(map(float, x.split()))[:3] for x in lines[end + 1:end + 1 + nv]])
    F = np.array([list(map(int, x.split()))[1:] for x in lines[end + 1 + nv:] if x.strip()])
    assert np.array_equal(F, t)
    tok = g["scn"].tobytes().decode().split()
    tr, sc, q = np.array(tok[2:5], float), float(tok[5]), np.array(tok[7:11], float)
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    np.testing.assert_allclose(v[:, :3], tr + sc * P @ R.T, rtol=0, atol=2e-6 * sc)
    fn = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    N = np.zeros_like(P)
    for k in range(3):
        np.add.at(N, F[:, k], fn)
    N = N @ R.T
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    np.testing.assert_allclose(v[:, 3:], N, rtol=0, atol=2e-5)


def test_malformed_ply_is_refused(tmp_path):
    """Header counts beyond what the file can hold, huge polygon sizes and non-numeric
    bodies fail with an error (no allocation from unchecked counts)."""
    head = "ply\nformat ascii 1.0\nelement vertex {nv}\nproperty float x\nproperty float y\nproperty float z\n" \
           "element face {nf}\nproperty list uchar int vertex_indices\nend_header\n"
    mesh = tmp_path / "m.scn"
    mesh.write_text("mesh bad.ply 0 0 0 1 q 0 0 0 1\n")
    cases = [head.format(nv=10 ** 15, nf=1) + "0 0 0\n",
             head.format(nv=3, nf=10 ** 15) + "0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n",
             head.format(nv=3, nf=1) + "0 0 0\n1 0 0\n0 1 0\n4000000000 0 1 2\n",
             head.format(nv=3, nf=1) + "0 0 0\n1 x 0\n0 1 0\n3 0 1 2\n"]
    for body in cases:
        (tmp_path / "bad.ply").write_text(body)
        with pytest.raises(rtx.RtxError):
            rtx.load_scene(mesh)


def _write_bvh(path, version=1, truncate=0):
    nodes = np.zeros(1, O.NODE48)
    nodes["is_leaf"] = 1
    nodes["b"] = 1
    tris = np.array([[0, 1, 2]], np.uint32)
    verts = np.zeros((3, 6), np.float32)
    blob = np.array([version, 1, 1, 3], np.uint32).tobytes() + nodes.tobytes() + tris.tobytes() + verts.tobytes()
    path.write_bytes(blob[: len(blob) - truncate])


def test_bvh_file_errors(tmp_path):
    with pytest.raises(rtx.RtxError, match="failed to open"):
        rtx.load_bvh(tmp_path / "none.bvh")
    _write_bvh(tmp_path / "v2.bvh", version=2)
    with pytest.raises(rtx.RtxError, match="version mismatch"):
        rtx.load_bvh(tmp_path / "v2.bvh")
    _write_bvh(tmp_path / "short.bvh", truncate=10)
    with pytest.raises(rtx.RtxError, match="payload"):
        rtx.load_bvh(tmp_path / "short.bvh")
    (tmp_path / "hdr.bvh").write_bytes(b"\x01\x00")
    with pytest.raises(rtx.RtxError, match="header"):
        rtx.load_bvh(tmp_path / "hdr.bvh")


def _closed_box():
    parts = []
    for axis in range(3):
        for sgn in (-1, 1):
            c = np.zeros(3)
            c[axis] = sgn
            u = np.zeros(3); u[(axis + 1) % 3] = 1
            w = np.zeros(3); w[(axis + 2) % 3] = 1
            pts = [c - u - w, c + u - w, c + u + w, c - u + w]
            parts.append(S.quad(*pts, -c))
    return S.merge(parts)


@pytest.mark.parametrize("mb", [0, 1, 4])
def test_oracle_white_furnace(mb):
    """Closed box, albedo a and emission E everywhere, no roulette: every path hits a wall
    at every bounce, so each pixel is E * sum_{k<=mb} a^k exactly."""
    v, t = _closed_box()
    o = O.build(v, t)
    a, E = 0.6, 0.5
    albedo = np.full((len(t), 3), a, np.float32)
    emission = np.full((len(t), 3), E, np.float32)
    cam = S.camera_lookat((0.1, 0.05, -0.2), (0.3, 0.2, 1.0), 60.0)
    img = O.render(o, cam, O.make_params(16, 12, 3, mb, 99, seed=1), albedo, emission)
    np.testing.assert_allclose(img, E * sum(a ** k for k in range(mb + 1)), rtol=2e-6)


def test_hdr_round_trip_and_flip(tmp_path):
    rng = np.random.default_rng(1)
    img = (rng.random((7, 9, 3)) * 4).astype(np.float32)
    img[0, 0] = 0
    rtx.write_hdr(tmp_path / "a.hdr", img, flip_y=False)
    back = rtx.read_hdr(tmp_path / "a.hdr")
    assert back.shape == img.shape
    # RGBE shares one exponent per pixel: error <= (largest channel) / 128, truncated
    assert (np.abs(back - img) <= img.max(-1, keepdims=True) / 128 + 1e-7).all()
    rtx.write_hdr(tmp_path / "b.hdr", img, flip_y=True)
    np.testing.assert_array_equal(rtx.read_hdr(tmp_path / "b.hdr"), back[::-1])


def test_hdr_reads_run_length_scanlines(tmp_path):
    """New-style RLE scanlines (channel planes, runs > 128 repeat the next byte)."""
    w, h = 10, 2
    body = b""
    for _ in range(h):
        body += bytes([2, 2, 0, w])
        body += bytes([128 + 10, 100])                    # R: run of 10
        body += bytes([10]) + bytes(range(10))             # G: 10 literals
        body += bytes([128 + 10, 0])                       # B: run of zeros
        body += bytes([128 + 10, 129])                     # E: exponent 1
    (tmp_path / "r.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 10\n" + body)
    img = rtx.read_hdr(tmp_path / "r.hdr")
    f = 2.0 ** (129 - 136)
    np.testing.assert_allclose(img[:, :, 0], 100 * f)
    np.testing.assert_allclose(img[1, :, 1], np.arange(10) * f)
    assert (img[:, :, 2] == 0).all()


def test_oracle_ibl_lookup():
    """A constant map equals the constant environment; a map whose rows differ picks the
    row of the ray's polar angle (row 0 = straight up, IBL::sample main.cpp:107-117)."""
    v = np.array([[0, 0, 10, 0, 0, -1], [1, 0, 10, 0, 0, -1], [0, 1, 10, 0, 0, -1]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    o = O.build(v, t)
    cam = S.camera_lookat((0, 0, 0), (0, 0, -1), 60.0)
    const = np.full((8, 16, 3), 0.25, np.float32)
    a = O.render(o, cam, O.make_params(12, 8, 2, 3, 3, seed=1, env_mode=2, ibl=const))
    b = O.render(o, cam, O.make_params(12, 8, 2, 3, 3, seed=1, env_mode=1, env_color=(0.25, 0.25, 0.25)))
    np.testing.assert_array_equal(a, b)
    rows = np.repeat(np.arange(4, dtype=np.float32)[:, None, None], 16, 1).repeat(3, 2)
    up = S.camera_lookat((0, 0, 0), (0, 1, 0.0001), 10.0)
    img = O.render(o, up, O.make_params(4, 4, 1, 0, 3, seed=1, env_mode=2, ibl=rows))
    assert (img == 0).all()                   # looking straight up: the top row
    down = S.camera_lookat((0, 0, 0), (0, -1, 0.0001), 10.0)
    img = O.render(o, down, O.make_params(4, 4, 1, 0, 3, seed=1, env_mode=2, ibl=rows))
    assert (img == 3).all()                   # straight down: the bottom row


def test_tuning_validation():
    """rtx_set_tuning accepts only launch shapes the kernels support; a refused call
    leaves the tuning unchanged."""
    saved = rtx.get_tuning()
    try:
        for bad in (dict(trace_loop=3), dict(trace_lds_depth=10), dict(shade_threads=96),
                    dict(batch_paths=512), dict(grid_blocks=0), dict(ray_sort=2), dict(wait_lanes=65),
                    dict(trace_grid_blocks=0), dict(render_lds_depth=4), dict(render_loop=3), dict(render_wait_lanes=99),
                    dict(refill_lanes=0), dict(refill_lanes=65), dict(trace_tree=2),
                    dict(leaf_coop=2), dict(render_leaf_coop=2)):
            with pytest.raises(rtx.RtxError):
                rtx.set_tuning(**bad)
            assert rtx.get_tuning() == saved, bad
        rtx.set_tuning(ray_sort=1, trace_lds_depth=12, trace_loop=0)
        assert rtx.get_tuning() == dict(saved, ray_sort=1, trace_lds_depth=12, trace_loop=0)
    finally:
        rtx.set_tuning(**saved)


@pytest.mark.parametrize("wh", [(1, 1), (7, 3), (8, 8), (9, 17), (33, 5), (1920, 1080), (17, 1), (1, 17)])
def test_tile_pixel_order_is_a_bijection(wh):
    """The renderer's path order within a sample (8 x 8 tiles along bands of 8 rows) visits
    every pixel once; a full tile is 64 consecutive slots covering an 8 x 8 block."""
    W, H = wh
    k = np.arange(W * H)
    p = rtx.tile_pixel(k, W, H)
    assert np.array_equal(np.sort(p), k)
    if W >= 8 and H >= 8:
        blk = p[:64]
        assert sorted(set(blk % W)) == list(range(8)) and sorted(set(blk // W)) == list(range(8))
