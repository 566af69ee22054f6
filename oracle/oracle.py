"""ctypes bindings for the CPU oracle (liboracle.so) and the compiled-reference
harness (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product (librtx) never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

NODE48 = np.dtype([("bmin", "<f4", 4), ("bmax", "<f4", 4), ("is_leaf", "u1"), ("pad0", "u1", 3),
                   ("a", "<u4"), ("b", "<u4"), ("pad1", "<u4")])
assert NODE48.itemsize == 48


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tan_half_fov_y", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("max_bounce", C.c_int),
                ("rr_bounce", C.c_int), ("seed", C.c_uint32), ("env_mode", C.c_int),
                ("env_color", C.c_float * 3), ("ibl", C.c_void_p), ("ibl_w", C.c_int), ("ibl_h", C.c_int),
                ("uv", C.c_void_p), ("tex", C.c_void_p), ("tex_w", C.c_int), ("tex_h", C.c_int),
                ("sun_dir", C.c_float * 3), ("sun_cos", C.c_float), ("sun_rad", C.c_float * 3)]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_ORC = None
_REF = None


def lib():
    global _ORC
    if _ORC is None:
        _ORC = _load(ORACLE_SO)
    return _ORC


def mutant_path(k):
    """Planted-mistake build k of the oracle (ORC_MUTANT in rtx_oracle.cpp)."""
    return os.path.join(HERE, "_build", f"liboracle_mut{k}.so")


class using:
    """Context manager: route this module's oracle calls through another build of the
    oracle library (the planted-mistake builds), restoring the real one afterwards."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _ORC
        self.saved = lib()
        _ORC = _load(self.path)
        return _ORC

    def __exit__(self, *exc):
        global _ORC
        _ORC = self.saved
        return False


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
    L = C.CDLL(path)
    vp, u32, sz = C.c_void_p, C.c_uint32, C.c_size_t
    L.orc_morton.argtypes = [vp, u32, vp, u32, vp, vp]
    L.orc_sort.argtypes = [vp, u32, vp]
    L.orc_build.argtypes = [vp, u32, vp, u32, C.POINTER(vp)]
    L.orc_tree_free.argtypes = [vp]
    for f in ("orc_tree_counts", "orc_tree_nodes", "orc_tree_tris", "orc_tree_prims",
              "orc_tree_sorted", "orc_tree_bfs_nodes"):
        getattr(L, f).argtypes = [vp, vp]
        getattr(L, f).restype = None
    L.orc_trace.argtypes = [vp, u32, vp, u32, vp, u32, vp, sz, vp, vp, vp, C.c_int, C.c_int]
    L.orc_sample.argtypes = [u32, u32, u32, u32]
    L.orc_sample.restype = C.c_float
    L.orc_sincos.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_atan2.argtypes = [C.c_float, C.c_float]
    L.orc_atan2.restype = C.c_float
    L.orc_render.argtypes = [vp, u32, vp, vp, u32, vp, u32, vp, vp, C.POINTER(Camera),
                             C.POINTER(RenderParams), vp, C.c_int]
    L.orc_render_loss_f64.argtypes = [vp, u32, vp, vp, u32, vp, u32, vp, vp, C.POINTER(Camera),
                                      C.POINTER(RenderParams), vp, C.c_int]
    L.orc_render_loss_f64.restype = C.c_double
    return L


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _REF
    if _REF is None:
        if not ref_available():
            raise RuntimeError(f"compiled reference not built: {REF_SO}")
        L = C.CDLL(REF_SO)
        vp, u32, sz = C.c_void_p, C.c_uint32, C.c_size_t
        L.ref_build.argtypes = [vp, u32, vp, u32]
        for f in ("ref_counts", "ref_nodes", "ref_bfs_nodes", "ref_tris", "ref_prims", "ref_sorted",
                  "ref_codes", "ref_scene_box"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = None
        L.ref_trace.argtypes = [vp, u32, vp, u32, vp, u32, vp, sz, vp, vp]
        L.ref_geo.argtypes = [C.c_int, vp, vp, vp, vp, vp]
        L.ref_clip.argtypes = [vp, vp, vp]
        _REF = L
    return _REF


def _prep(verts6, tris):
    v = np.ascontiguousarray(verts6, dtype=np.float32).reshape(-1, 6)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    return v, t


class Tree:
    """A built tree in the reference's on-disk layout (BVHNode / Triangle arrays)."""

    def __init__(self, nodes, tris, prims, verts6, sorted_tris=None, bfs_nodes=None, info=None):
        self.nodes, self.tris, self.prims, self.verts = nodes, tris, prims, verts6
        self.sorted, self.bfs_nodes, self.info = sorted_tris, bfs_nodes, info or {}


def morton(verts6, tris):
    v, t = _prep(verts6, tris)
    codes = np.zeros(len(t), np.uint64)
    box = np.zeros(6, np.float32)
    rc = lib().orc_morton(_p(v), len(v), _p(t), len(t), _p(codes), _p(box))
    if rc:
        raise ValueError(f"orc_morton failed ({rc})")
    return codes, box


def sort(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint64)
    out = np.zeros(len(codes), np.uint32)
    lib().orc_sort(_p(codes), len(codes), _p(out))
    return out


def build(verts6, tris):
    v, t = _prep(verts6, tris)
    L = lib()
    h = C.c_void_p()
    rc = L.orc_build(_p(v), len(v), _p(t), len(t), C.byref(h))
    if rc:
        raise ValueError(f"orc_build failed ({rc})")
    try:
        c = np.zeros(7, np.uint32)
        L.orc_tree_counts(h, _p(c))
        nodes = np.zeros(c[0], NODE48)
        L.orc_tree_nodes(h, _p(nodes))
        rt = np.zeros((c[1], 3), np.uint32)
        L.orc_tree_tris(h, _p(rt))
        prims = np.zeros(c[1], np.uint32)
        L.orc_tree_prims(h, _p(prims))
        srt = np.zeros(len(t), np.uint32)
        L.orc_tree_sorted(h, _p(srt))
        info = dict(nodes=int(c[0]), tris=int(c[1]), tindex_size=int(c[2]), sorted_size=int(c[3]),
                    depth=int(c[4]), spatial=int(c[5]), budget_rejected=int(c[6]))
        bfs = np.zeros(int(c[0]), NODE48)
        L.orc_tree_bfs_nodes(h, _p(bfs))
    finally:
        L.orc_tree_free(h)
    return Tree(nodes, rt, prims, v, srt, bfs, info)


def trace(tree, rays, any_hit=False, counts=False, normals=False, nthreads=1):
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    n = len(rays)
    hits = np.zeros((n, 4), np.float32)
    cnt = np.zeros((n, 2), np.uint32) if counts else None
    nrm = np.zeros((n, 3), np.float32) if normals else None
    rc = lib().orc_trace(_p(tree.nodes), len(tree.nodes), _p(tree.tris), len(tree.tris), _p(tree.verts),
                         len(tree.verts), _p(rays), n, _p(hits), _p(nrm), _p(cnt), int(any_hit), nthreads)
    if rc:
        raise ValueError(f"orc_trace failed ({rc})")
    out = {"t": hits[:, 0].copy(), "u": hits[:, 1].copy(), "v": hits[:, 2].copy(),
           "tri": hits[:, 3].view(np.uint32).copy()}
    if counts:
        out["counts"] = cnt
    if normals:
        out["normal"] = nrm
    return out


def make_camera(cam):
    c = Camera()
    for k in ("pos", "fwd", "right", "up"):
        getattr(c, k)[:] = [float(x) for x in cam[k]]
    c.tan_half_fov_y = float(cam["tan_half_fov_y"])
    return c


def make_params(width, height, spp, max_bounce, rr_bounce=3, seed=0, env_mode=0, env_color=(0, 0, 0), ibl=None,
                uv=None, tex=None, sun=None):
    """ibl: (h, w, 3) float32 array, top row first (env_mode 2); uv: (nverts, 2) and tex:
    (h, w, 3) for texture albedo; sun: (direction, cos of the cone, radiance) for env_mode 3.
    Arrays are kept alive on the params."""
    p = RenderParams()
    p.width, p.height, p.spp, p.max_bounce, p.rr_bounce, p.seed = width, height, spp, max_bounce, rr_bounce, seed
    p.env_mode = env_mode
    p.env_color[:] = [float(x) for x in env_color]
    p._keep = []
    if ibl is not None:
        arr = np.ascontiguousarray(ibl, np.float32)
        p._keep.append(arr)
        p.ibl = arr.ctypes.data
        p.ibl_h, p.ibl_w = arr.shape[:2]
    if tex is not None:
        set_texture(p, uv, tex)
    if sun is not None:
        d, c, rad = sun
        p.sun_dir[:] = [float(x) for x in d]
        p.sun_cos = float(c)
        p.sun_rad[:] = [float(x) for x in rad]
    return p


def set_texture(p, uv, tex):
    """Point the params at a (possibly new) texture; uv (nverts, 2), tex (h, w, 3)."""
    u = np.ascontiguousarray(uv, np.float32)
    t = np.ascontiguousarray(tex, np.float32)
    p._keep += [u, t]
    p.uv, p.tex = u.ctypes.data, t.ctypes.data
    p.tex_h, p.tex_w = t.shape[:2]


def render(tree, cam, params, albedo=None, emission=None, nthreads=1):
    img = np.zeros((params.height, params.width, 3), np.float32)
    a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
    e = None if emission is None else np.ascontiguousarray(emission, np.float32)
    rc = lib().orc_render(_p(tree.nodes), len(tree.nodes), _p(tree.tris), _p(tree.prims), len(tree.tris),
                          _p(tree.verts), len(tree.verts), _p(a), _p(e), C.byref(make_camera(cam)),
                          C.byref(params), _p(img), nthreads)
    if rc:
        raise ValueError(f"orc_render failed ({rc})")
    return img


def render_loss_f64(tree, cam, params, albedo, emission, target, nthreads=1):
    a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
    e = None if emission is None else np.ascontiguousarray(emission, np.float32)
    tg = np.ascontiguousarray(target, np.float32)
    return lib().orc_render_loss_f64(_p(tree.nodes), len(tree.nodes), _p(tree.tris), _p(tree.prims),
                                     len(tree.tris), _p(tree.verts), len(tree.verts), _p(a), _p(e),
                                     C.byref(make_camera(cam)), C.byref(params), _p(tg), nthreads)


def sample(seed, pixel, s, dim):
    return lib().orc_sample(seed, pixel, s, dim)


def atan2(y, x):
    return lib().orc_atan2(float(y), float(x))


def sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


# ---- compiled reference -------------------------------------------------------------

def ref_build(verts6, tris):
    v, t = _prep(verts6, tris)
    R = ref()
    rc = R.ref_build(_p(v), len(v), _p(t), len(t))
    if rc:
        raise ValueError(f"ref_build failed ({rc})")
    c = np.zeros(4, np.uint32)
    R.ref_counts(_p(c))
    nodes = np.zeros(c[0], NODE48)
    R.ref_nodes(_p(nodes))
    rt = np.zeros((c[1], 3), np.uint32)
    R.ref_tris(_p(rt))
    prims = np.zeros(c[1], np.uint32)
    R.ref_prims(_p(prims))
    srt = np.zeros(len(t), np.uint32)
    R.ref_sorted(_p(srt))
    codes = np.zeros(len(t), np.uint64)
    R.ref_codes(_p(codes))
    box = np.zeros(6, np.float32)
    R.ref_scene_box(_p(box))
    bfs = np.zeros(c[0], NODE48)
    R.ref_bfs_nodes(_p(bfs))
    tr = Tree(nodes, rt, prims, v, srt, bfs,
              dict(nodes=int(c[0]), tris=int(c[1]), tindex_size=int(c[2]), sorted_size=int(c[3])))
    tr.codes, tr.scene_box = codes, box
    return tr


def ref_trace(tree, rays, normals=False):
    """bvh::trace over `rays`; tree=None reuses the tree loaded by the previous call."""
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    n = len(rays)
    hits = np.zeros((n, 4), np.float32)
    nrm = np.zeros((n, 3), np.float32) if normals else None
    if tree is None:
        ref().ref_trace(None, 0, None, 0, None, 0, _p(rays), n, _p(hits), _p(nrm))
    else:
        ref().ref_trace(_p(tree.nodes), len(tree.nodes), _p(tree.tris), len(tree.tris), _p(tree.verts),
                        len(tree.verts), _p(rays), n, _p(hits), _p(nrm))
    out = {"t": hits[:, 0].copy(), "u": hits[:, 1].copy(), "v": hits[:, 2].copy(),
           "tri": hits[:, 3].view(np.uint32).copy()}
    if normals:
        out["normal"] = nrm
    return out


def ref_geo_sphere(center, radius, rings=16, sectors=32):
    args = np.array([*center, radius, rings, sectors, 0, 0, 0, 0], np.float32)
    return _ref_geo(0, args)


def ref_geo_box(center, half, quat_xyzw=(0, 0, 0, 1)):
    args = np.array([*center, *half, *quat_xyzw], np.float32)
    return _ref_geo(1, args)


def _ref_geo(kind, args):
    R = ref()
    nv, nt = C.c_uint32(), C.c_uint32()
    R.ref_geo(kind, _p(args), None, C.byref(nv), None, C.byref(nt))
    v = np.zeros((nv.value, 6), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    R.ref_geo(kind, _p(args), _p(v), C.byref(nv), _p(t), C.byref(nt))
    return v, t


def ref_clip(v9, region6):
    v = np.ascontiguousarray(v9, np.float32)
    r = np.ascontiguousarray(region6, np.float32)
    out = np.zeros(6, np.float32)
    ok = ref().ref_clip(_p(v), _p(r), _p(out))
    return bool(ok), out
