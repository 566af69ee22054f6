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
    return out


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
