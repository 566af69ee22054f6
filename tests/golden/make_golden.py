"""Writes tests/golden/reference_*.npz: outputs of the compiled reference (oracle/_ref,
built from the reference's own sources by oracle/Makefile) on small seeded scenes, so
the oracle stays pinned where the reference checkout is absent.

reference_clip.npz holds triangle/box pairs and clipTriangleToAABB's boxes (AABB.cpp:86-171).

Each scene file holds the inputs (vertices, triangles, rays) and the reference's outputs:
Morton codes and scene box, sort order, builder-order and DFS-order nodes (raw 48-byte
BVHNode records), reordered triangles and original ids, s_triIndex / s_sortedTris sizes,
and closest-hit t/u/v/triIndex/normal for a random ray set and an edge-case ray set.

    python tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "ray-tracer_amd"), os.path.dirname(HERE)]

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402
from scenes_util import edge_rays, scene_suite  # noqa: E402
from rtx import scenes as S  # noqa: E402

SCENES = ["rand17", "rand1000", "cornell", "test_traverse", "straddle", "dupcodes", "degenerate", "grid"]


def rays_for(v, n, seed):
    lo, hi = v[:, :3].min(0) - 0.1, v[:, :3].max(0) + 0.1
    r = S.incoherent_rays(n, seed=seed)
    r[:, :3] = lo + (hi - lo) * r[:, :3]
    return r


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (run make -C oracle where the reference exists)")
    suite = scene_suite(O.ref_geo_sphere, O.ref_geo_box)
    for name in SCENES:
        v, t = suite[name]
        r = O.ref_build(v, t)
        rays = rays_for(v, 1000, seed=len(t))
        edge = edge_rays(v, 1000, seed=3)
        out = {"verts": np.ascontiguousarray(v, np.float32), "tris": np.ascontiguousarray(t, np.uint32),
               "codes": r.codes, "scene_box": r.scene_box, "sorted": r.sorted,
               "nodes": r.nodes.view(np.uint8).reshape(-1, 48), "bfs_nodes": r.bfs_nodes.view(np.uint8).reshape(-1, 48),
               "rtris": r.tris, "prims": r.prims,
               "sizes": np.array([r.info["tindex_size"], r.info["sorted_size"]], np.uint32)}
        for key, rs in (("rays", rays), ("edge", edge)):
            h = O.ref_trace(r, rs, normals=True)
            out[key] = rs
            out[key + "_hit"] = np.stack([h["t"], h["u"], h["v"], h["tri"].view(np.float32)], 1)
            out[key + "_normal"] = h["normal"]
        path = os.path.join(HERE, f"reference_{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{path}: {len(t)} triangles, {len(r.nodes)} nodes, {os.path.getsize(path)} bytes")


def clip_cases(n=6000, seed=11):
    """Triangle/box pairs for the builder's clip: random and grid-valued vertices (ties on
    planes), signed zeros, degenerate triangles; the box is the triangle's own extent on two
    axes (sometimes cut) and a random slab on the third."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    grid = (rng.integers(0, 10, (n, 9)) * 0.125 - 0.5).astype(np.float32)
    v[1::4] = grid[1::4]
    z = rng.integers(0, 10, (n, 9)) == 0
    zm = np.zeros((n, 9), bool)
    zm[2::4] = z[2::4]
    v[zm] = np.where(rng.integers(0, 2, zm.sum()) == 1, np.float32(0.0), np.float32(-0.0))
    v[3::4, 3:6] = v[3::4, 0:3]
    axis = np.arange(n) % 3
    box = np.zeros((n, 6), np.float32)
    for i in range(n):
        p = v[i].reshape(3, 3)
        lo, hi = p.min(0), p.max(0)
        for b in range(3):
            if b == axis[i]:
                s, e = np.sort(lo[b] + (hi[b] - lo[b]) * rng.uniform(0, 1, 2).astype(np.float32))
                if i % 4 == 1:
                    s = np.float32(rng.integers(0, 10) * 0.125 - 0.5)
                    e = np.float32(s + 0.125 * (1 + rng.integers(0, 3)))
                box[i, b], box[i, 3 + b] = s, e
            elif rng.integers(0, 10) == 0:
                box[i, b], box[i, 3 + b] = lo[b] + np.float32(0.25) * (hi[b] - lo[b]), hi[b]
            else:
                box[i, b], box[i, 3 + b] = lo[b], hi[b]
    out = np.zeros((n, 6), np.float32)
    valid = np.zeros(n, bool)
    for i in range(n):
        valid[i], out[i] = O.ref_clip(v[i], box[i])
    path = os.path.join(HERE, "reference_clip.npz")
    np.savez_compressed(path, verts=v, region=box, axis=axis.astype(np.int32), box=out, valid=valid)
    print(f"{path}: {n} clip cases, {os.path.getsize(path)} bytes")


BUNNY = "/root/reference/bunny.ply"
# translate (0.5, -0.25, 1), uniform scale 6, rotation by a unit quaternion (x, y, z, w)
BUNNY_SCN = "mesh bunny_head.ply 0.5 -0.25 1.0 6.0 q 0.1 0.2 0.3 0.9273618495495703\n"


def bunny_head(keep=0.12):
    """ASCII PLY text of the top `keep` fraction (by height) of the reference's bunny.ply:
    every face whose three vertices lie above that height, with its vertices renumbered in
    file order and the file's header layout (x y z confidence intensity, uchar/int lists)."""
    with open(BUNNY) as f:
        lines = f.read().split("\n")
    end = lines.index("end_header")
    nv = int(next(x for x in lines if x.startswith("element vertex")).split()[2])
    nf = int(next(x for x in lines if x.startswith("element face")).split()[2])
    vals = np.array([list(map(float, x.split())) for x in lines[end + 1:end + 1 + nv]])
    faces = np.array([list(map(int, x.split())) for x in lines[end + 1 + nv:end + 1 + nv + nf]])
    assert (faces[:, 0] == 3).all()
    cut = np.quantile(vals[:, 1], 1.0 - keep)
    f = faces[:, 1:]
    f = f[(vals[f, 1] > cut).all(1)]
    used = np.unique(f)
    remap = -np.ones(nv, np.int64)
    remap[used] = np.arange(len(used))
    vtext = [lines[end + 1 + i] for i in used]
    ftext = ["3 %d %d %d" % tuple(remap[x]) for x in f]
    header = ["ply", "format ascii 1.0", f"element vertex {len(used)}", "property float x", "property float y",
              "property float z", "property float confidence", "property float intensity",
              f"element face {len(f)}", "property list uchar int vertex_indices", "end_header"]
    return "\n".join(header + vtext + ftext) + "\n"


def bunny_golden():
    """reference_bunny_head.npz: a mesh read from PLY by rtx_scene_load (the .scn `mesh` line,
    bvh_IO.cpp:195-231, in place of assimp's loadMesh), then built and traced by the compiled
    reference. Holds the PLY and .scn text so the tests can load the same files."""
    import tempfile
    import rtx
    ply = bunny_head()
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "bunny_head.ply"), "w") as f:
            f.write(ply)
        with open(os.path.join(d, "head.scn"), "w") as f:
            f.write(BUNNY_SCN)
        v, t = rtx.load_scene(os.path.join(d, "head.scn"))
    r = O.ref_build(v, t)
    out = {"ply": np.frombuffer(ply.encode(), np.uint8), "scn": np.frombuffer(BUNNY_SCN.encode(), np.uint8),
           "verts": v, "tris": t, "codes": r.codes, "scene_box": r.scene_box, "sorted": r.sorted,
           "nodes": r.nodes.view(np.uint8).reshape(-1, 48), "bfs_nodes": r.bfs_nodes.view(np.uint8).reshape(-1, 48),
           "rtris": r.tris, "prims": r.prims,
           "sizes": np.array([r.info["tindex_size"], r.info["sorted_size"]], np.uint32)}
    for key, rs in (("rays", rays_for(v, 1000, seed=len(t))), ("edge", edge_rays(v, 1000, seed=5))):
        h = O.ref_trace(r, rs, normals=True)
        out[key] = rs
        out[key + "_hit"] = np.stack([h["t"], h["u"], h["v"], h["tri"].view(np.float32)], 1)
        out[key + "_normal"] = h["normal"]
    path = os.path.join(HERE, "reference_bunny_head.npz")
    np.savez_compressed(path, **out)
    print(f"{path}: {len(t)} triangles, {len(r.nodes)} nodes, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
    clip_cases()
    bunny_golden()
