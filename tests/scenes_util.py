"""Shared test scenes (host arrays) and tree comparison helpers."""
import numpy as np

from rtx import scenes as S


def straddle_scene(n=3000, seed=7):
    """Large overlapping triangles: most cross any split plane, so spatial splits are
    chosen often and the 1.5 N reference budget runs out."""
    return S.random_triangles(n, seed=seed, mu=-1.0, sigma=0.3)


def reference_test_scene(geo_sphere, geo_box):
    """The scene of the reference's tests/test_traverse.cpp:73-76."""
    return S.merge([geo_box((0, 0, 5), (1, 1, 1)), geo_box((0, 0, 12), (1, 1, 1)),
                    geo_sphere((0, 5, 0), 2.0), geo_box((0, -3, 0), (20, 0.1, 20))])


def scene_suite(geo_sphere, geo_box):
    out = {}
    for n in (1, 5, 17, 65, 100, 1000, 20000):
        out[f"rand{n}"] = S.random_triangles(n, seed=n)
    v, t, _, _, _ = S.cornell_box()
    out["cornell"] = (v, t)
    out["test_traverse"] = reference_test_scene(geo_sphere, geo_box)
    out["straddle"] = straddle_scene()
    # identical Morton codes: many copies of one tiny triangle plus a few others
    base, bt = S.random_triangles(300, seed=11, mu=-9.0)
    base[:600, :3] = base[0:3, :3].reshape(1, 3, 3).repeat(200, 0).reshape(600, 3)
    out["dupcodes"] = (base, bt)
    out["degenerate"] = degenerate_scene()
    out["grid"] = grid_scene()
    return out


def degenerate_scene():
    """Random soup where every tenth triangle of each kind is degenerate: all three
    vertices equal, collinear vertices, two equal vertices, or an exact copy of its
    predecessor (identical boxes, centroids and Morton codes)."""
    v, t = S.random_triangles(3000, seed=31)
    p = v.reshape(-1, 3, 6)
    p[0::10, 1, :3] = p[0::10, 0, :3]
    p[0::10, 2, :3] = p[0::10, 0, :3]
    p[3::10, 2, :3] = p[3::10, 0, :3] + 2.0 * (p[3::10, 1, :3] - p[3::10, 0, :3])
    p[5::10, 1, :3] = p[5::10, 0, :3]
    p[7::10] = p[6::10][: len(p[7::10])]
    return p.reshape(-1, 6).copy(), t


def grid_scene(m=12):
    """Unit squares (two triangles each) on integer lattice planes of all three
    orientations: node boxes share planes exactly and rays through grid lines hit shared
    edges at exactly equal t."""
    quads = []
    for axis in range(3):
        for level in (0, 3, 7):
            for i in range(m):
                for j in range(m):
                    if (i * 7 + j * 3 + level + axis) % 4 == 0:
                        continue
                    a, b = (axis + 1) % 3, (axis + 2) % 3
                    c = []
                    for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        q = [0.0, 0.0, 0.0]
                        q[axis] = float(level)
                        q[a] = float(i + di)
                        q[b] = float(j + dj)
                        c.append(q)
                    quads.append(c)
    q = np.array(quads, np.float32)                       # (Q, 4, 3)
    tri_pos = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]], q[:, [0, 2, 1]], q[:, [0, 3, 2]]], 0)
    n = len(tri_pos)
    nrm = np.cross(tri_pos[:, 1] - tri_pos[:, 0], tri_pos[:, 2] - tri_pos[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    v = np.empty((n, 3, 6), np.float32)
    v[:, :, :3] = tri_pos
    v[:, :, 3:] = nrm[:, None, :]
    return v.reshape(-1, 6), np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def edge_rays(v, n, seed):
    """Rays that take every branch of the slab test's NaN handling, mixed in one set:
    random rays; one or two zero direction components (infinite inverse); -0.0 components;
    denormal components (the inverse overflows to inf); origins exactly on vertex
    coordinates of the scene (box planes) with a zero component on that axis, which makes
    (plane - origin) * inverse = 0 * inf = NaN; tmin/tmax windows."""
    rng = np.random.default_rng(seed)
    lo, hi = v[:, :3].min(0), v[:, :3].max(0)
    r = S.incoherent_rays(n, seed=seed)
    r[:, :3] = lo - 0.05 + (hi - lo + 0.1) * r[:, :3]
    kind = np.arange(n) % 8
    ax = rng.integers(0, 3, n)
    idx = np.arange(n)
    sel = kind == 1
    r[idx[sel], 4 + ax[sel]] = 0.0
    sel = kind == 2
    r[idx[sel], 4 + ax[sel]] = 0.0
    r[idx[sel], 4 + (ax[sel] + 1) % 3] = 0.0
    sel = kind == 3
    r[idx[sel], 4 + ax[sel]] = -0.0
    sel = kind == 4
    r[idx[sel], 4 + ax[sel]] = np.float32(1e-40) * np.where(rng.random(sel.sum()) < 0.5, 1, -1)
    sel = np.nonzero((kind == 5) | (kind == 6))[0]
    pick = rng.integers(0, len(v), len(sel))
    r[sel, ax[sel]] = v[pick, ax[sel]]
    r[sel, 4 + ax[sel]] = 0.0
    sel = kind == 7
    r[sel, 3] = rng.random(sel.sum()).astype(np.float32) * 0.3
    r[sel, 7] = r[sel, 3] + rng.random(sel.sum()).astype(np.float32) * 0.5
    bad = np.all(r[:, 4:7] == 0, axis=1)
    r[bad, 4] = 1.0
    return r


def assert_same_nodes(a, b, what=""):
    for f in ("bmin", "bmax"):
        x = np.ascontiguousarray(a[f][:, :3]).view(np.uint32)
        y = np.ascontiguousarray(b[f][:, :3]).view(np.uint32)
        if not np.array_equal(x, y):
            i = np.nonzero((x != y).any(1))[0][0]
            raise AssertionError(f"{what}: node {i} {f} differs: {a[f][i]} vs {b[f][i]}")
    for f in ("is_leaf", "a", "b"):
        if not np.array_equal(a[f], b[f]):
            i = np.nonzero(a[f] != b[f])[0][0]
            raise AssertionError(f"{what}: node {i} field {f} differs: {a[f][i]} vs {b[f][i]}")
