"""Python host side of librtx (MI355X BVH build, traversal and path tracing).

Mirrors the reference API of minsuk0519/ray-tracer for the hot path:
    bake_bvh(scene_path)   ~ bvh::bakeBVH   (src/bvh/bvh_bake.cpp:180-194)
    load_bvh(path)         ~ bvh::loadBVH   (src/bvh/bvh_traverse.cpp:182-235)
    BVH.trace(rays)        ~ bvh::trace     (src/bvh/bvh_traverse.cpp:237-246), batched
Everything runs through the C ABI of librtx.so (include/rtx.h) on the GPU; there is no
CPU fallback. Importing works without a GPU; calls that need one raise RtxError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(HERE, "..", "lib", "librtx.so"))

NODE48 = np.dtype([("bmin", "<f4", 4), ("bmax", "<f4", 4), ("is_leaf", "u1"), ("pad0", "u1", 3),
                   ("a", "<u4"), ("b", "<u4"), ("pad1", "<u4")])

# Every symbol include/rtx.h declares.
EXPORTS = ["rtx_last_error", "rtx_device_count", "rtx_build", "rtx_build_device", "rtx_bvh_free",
           "rtx_bvh_get_info", "rtx_bvh_export", "rtx_bvh_morton", "rtx_bvh_export_bfs", "rtx_debug_clip_slab", "rtx_debug_recip_check", "rtx_debug_scratch_peak",
           "rtx_trace_closest",
           "rtx_trace_closest_noalloc", "rtx_trace_any", "rtx_trace_counts", "rtx_trace_closest_host", "rtx_save", "rtx_load",
           "rtx_scene_load", "rtx_free_host", "rtx_bake_scene", "rtx_geo_sphere", "rtx_geo_box", "rtx_render",
           "rtx_render_primal", "rtx_render_adjoint", "rtx_render_adjoint_replay", "rtx_render_batched", "rtx_batch_create", "rtx_batch_launch", "rtx_batch_free",
           "rtx_set_tuning", "rtx_get_tuning", "rtx_hdr_write", "rtx_hdr_read", "rtx_profile_begin",
           "rtx_profile_end"]


class RtxError(RuntimeError):
    pass


class BvhInfo(C.Structure):
    _fields_ = [("node_count", C.c_uint32), ("tri_count", C.c_uint32), ("vert_count", C.c_uint32),
                ("input_tri_count", C.c_uint32), ("internal_count", C.c_uint32), ("leaf_count", C.c_uint32),
                ("max_depth", C.c_uint32), ("tri_index_size", C.c_uint32), ("sorted_tris_size", C.c_uint32),
                ("levels", C.c_uint32), ("scene_min", C.c_float * 3), ("scene_max", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tan_half_fov_y", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("max_bounce", C.c_int),
                ("rr_bounce", C.c_int), ("seed", C.c_uint32), ("env_mode", C.c_int),
                ("env_color", C.c_float * 3), ("ibl_d", C.c_void_p), ("ibl_width", C.c_int), ("ibl_height", C.c_int),
                ("uv_d", C.c_void_p), ("tex_d", C.c_void_p), ("tex_width", C.c_int), ("tex_height", C.c_int),
                ("sun_dir", C.c_float * 3), ("sun_cos", C.c_float), ("sun_radiance", C.c_float * 3)]


class Tuning(C.Structure):
    _fields_ = [("batch_paths", C.c_uint32), ("grid_blocks", C.c_uint32), ("shade_threads", C.c_uint32),
                ("trace_lds_depth", C.c_uint32), ("trace_loop", C.c_uint32), ("ray_sort", C.c_uint32),
                ("wait_lanes", C.c_uint32), ("trace_grid_blocks", C.c_uint32), ("render_lds_depth", C.c_uint32),
                ("render_loop", C.c_uint32), ("render_wait_lanes", C.c_uint32),
                ("leaf_coop", C.c_uint32), ("render_leaf_coop", C.c_uint32), ("refill_lanes", C.c_uint32),
                ("trace_tree", C.c_uint32)]


_LIB = None


def lib():
    """Load librtx.so; raise if it has not been built (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RtxError(f"librtx.so not built at {LIB_PATH}; run ray-tracer_amd/build.py")
        # torch ships its own HIP/HSA runtime and asks for it under unversioned names. Load
        # torch first so librtx's libamdhip64.so.7 / libhsa-runtime64.so.1 resolve (by
        # SONAME) to that same runtime: one HIP runtime per process, shared device memory.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, u32, sz, i32 = C.c_void_p, C.c_uint32, C.c_size_t, C.c_int
        L.rtx_last_error.restype = C.c_char_p
        L.rtx_build.argtypes = [vp, u32, vp, u32, C.POINTER(vp)]
        L.rtx_build_device.argtypes = [vp, u32, vp, u32, vp, C.POINTER(vp)]
        L.rtx_bvh_free.argtypes = [vp]
        L.rtx_bvh_free.restype = None
        L.rtx_bvh_get_info.argtypes = [vp, C.POINTER(BvhInfo)]
        L.rtx_bvh_export.argtypes = [vp, vp, vp, vp]
        L.rtx_bvh_morton.argtypes = [vp, vp, vp]
        L.rtx_debug_clip_slab.argtypes = [vp, vp, C.c_int, vp]
        L.rtx_debug_recip_check.argtypes = [C.c_uint64, C.c_uint64, vp, vp]
        L.rtx_debug_scratch_peak.argtypes = [vp, C.c_int]
        L.rtx_bvh_export_bfs.argtypes = [vp, vp, C.POINTER(u32)]
        L.rtx_trace_closest.argtypes = [vp, vp, sz, vp, vp, vp]
        L.rtx_trace_closest_noalloc.argtypes = [vp, vp, sz, vp, vp, vp]
        L.rtx_trace_any.argtypes = [vp, vp, sz, vp, vp]
        L.rtx_trace_counts.argtypes = [vp, vp, sz, vp, vp]
        L.rtx_trace_closest_host.argtypes = [vp, vp, sz, vp, vp]
        L.rtx_save.argtypes = [vp, C.c_char_p]
        L.rtx_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.rtx_scene_load.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(u32),
                                     C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(u32)]
        L.rtx_free_host.argtypes = [vp]
        L.rtx_free_host.restype = None
        L.rtx_bake_scene.argtypes = [C.c_char_p]
        L.rtx_geo_sphere.argtypes = [vp, C.c_float, i32, i32, vp, C.POINTER(u32), vp, C.POINTER(u32)]
        L.rtx_geo_box.argtypes = [vp, vp, vp, vp, C.POINTER(u32), vp, C.POINTER(u32)]
        L.rtx_render.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp]
        L.rtx_render_primal.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp, vp]
        L.rtx_render_adjoint.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp, vp, vp]
        L.rtx_render_adjoint_replay.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderParams), vp, vp, vp, vp, vp, vp,
                                                vp]
        L.rtx_render_batched.argtypes = [vp, i32, vp, vp, vp, vp, vp, i32, vp]
        L.rtx_batch_create.argtypes = [vp, i32, vp, vp, vp, vp, vp, i32, C.POINTER(vp)]
        L.rtx_batch_launch.argtypes = [vp, vp]
        L.rtx_batch_free.argtypes = [vp]
        L.rtx_batch_free.restype = None
        L.rtx_hdr_write.argtypes = [C.c_char_p, i32, i32, vp, i32]
        L.rtx_hdr_read.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(C.POINTER(C.c_float))]
        L.rtx_set_tuning.argtypes = [C.POINTER(Tuning)]
        L.rtx_get_tuning.argtypes = [C.POINTER(Tuning)]
        L.rtx_profile_begin.argtypes = []
        L.rtx_profile_end.argtypes = [C.POINTER(C.c_double), C.POINTER(u32)]
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != 0:
        msg = lib().rtx_last_error()
        raise RtxError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _dptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


_DTYPES = {"f32": ("float32",), "u32": ("int32", "uint32"), "u8": ("uint8", "bool")}


def _need(t, what, kind="f32", numel=None, optional=False):
    """Check a tensor handed to the library: a contiguous CUDA tensor of the given 4-byte
    float / 4-byte integer / byte kind holding at least `numel` elements. The kernels index
    these buffers by ray, pixel or triangle count, so anything shorter would be read or
    written out of bounds."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{what} is required")
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{what} must be a CUDA tensor")
    if str(t.dtype).replace("torch.", "") not in _DTYPES[kind]:
        raise ValueError(f"{what} must have dtype {' or '.join(_DTYPES[kind])}, not {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if numel is not None and t.numel() < numel:
        raise ValueError(f"{what} holds {t.numel()} elements, needs {numel}")
    return t


def _stream_ptr(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class BVH:
    """A tree resident in HBM (owned by librtx)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.rtx_bvh_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = BvhInfo()
        _check(lib().rtx_bvh_get_info(self._h, C.byref(inf)), "rtx_bvh_get_info")
        d = {k: getattr(inf, k) for k, _ in BvhInfo._fields_ if k not in ("scene_min", "scene_max")}
        d["scene_min"] = list(inf.scene_min)
        d["scene_max"] = list(inf.scene_max)
        return d

    def export(self):
        """(nodes [reference BVHNode layout, DFS order], reordered triangles, original ids)."""
        inf = self.info()
        nodes = np.zeros(inf["node_count"], NODE48)
        tris = np.zeros((inf["tri_count"], 3), np.uint32)
        prims = np.zeros(inf["tri_count"], np.uint32)
        _check(lib().rtx_bvh_export(self._h, _p(nodes), _p(tris), _p(prims)), "rtx_bvh_export")
        return nodes, tris, prims

    def export_bfs(self):
        n = C.c_uint32()
        _check(lib().rtx_bvh_export_bfs(self._h, None, C.byref(n)), "rtx_bvh_export_bfs")
        nodes = np.zeros(n.value, NODE48)
        _check(lib().rtx_bvh_export_bfs(self._h, _p(nodes), C.byref(n)), "rtx_bvh_export_bfs")
        return nodes

    def morton(self):
        n = self.info()["input_tri_count"]
        codes = np.zeros(n, np.uint64)
        srt = np.zeros(n, np.uint32)
        _check(lib().rtx_bvh_morton(self._h, _p(codes), _p(srt)), "rtx_bvh_morton")
        return codes, srt

    def save(self, path):
        _check(lib().rtx_save(self._h, str(path).encode()), "rtx_save")

    # ---- tracing --------------------------------------------------------------
    def trace(self, rays, normals=False):
        """Closest hit for host rays (N, 8) float32. Returns dict t/u/v/tri[/normal]."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        hits = np.zeros((n, 4), np.float32)
        nrm = np.zeros((n, 3), np.float32) if normals else None
        _check(lib().rtx_trace_closest_host(self._h, _p(rays), n, _p(hits), _p(nrm)), "rtx_trace_closest_host")
        out = {"t": hits[:, 0].copy(), "u": hits[:, 1].copy(), "v": hits[:, 2].copy(),
               "tri": hits[:, 3].view(np.uint32).copy()}
        if normals:
            out["normal"] = nrm
        return out

    @staticmethod
    def _rays(rays_d):
        _need(rays_d, "rays_d")
        if rays_d.dim() != 2 or rays_d.shape[1] != 8:
            raise ValueError("rays_d must have shape (N, 8)")
        return rays_d.shape[0]

    def trace_closest_device(self, rays_d, hits_d, normals_d=None, stream=None):
        """Device-resident closest hit: rays_d (N, 8) float32, hits_d (N, 4) 16-byte records
        (float32, or int32 to read the triangle index), normals_d optional (N, 3) float32."""
        n = self._rays(rays_d)
        if hits_d is None or str(hits_d.dtype) not in ("torch.float32", "torch.int32"):
            raise ValueError("hits_d must be a float32 or int32 CUDA tensor")
        _need(hits_d, "hits_d", "f32" if hits_d.dtype.is_floating_point else "u32", 4 * n)
        _need(normals_d, "normals_d", "f32", 3 * n, optional=True)
        _check(lib().rtx_trace_closest(self._h, _dptr(rays_d), n, _dptr(hits_d), _dptr(normals_d),
                                       _stream_ptr(stream)), "rtx_trace_closest")

    def trace_any_device(self, rays_d, occl_d, stream=None):
        n = self._rays(rays_d)
        _need(occl_d, "occl_d", "u8", n)
        _check(lib().rtx_trace_any(self._h, _dptr(rays_d), n, _dptr(occl_d), _stream_ptr(stream)),
               "rtx_trace_any")

    def trace_counts_device(self, rays_d, counts_d, stream=None):
        n = self._rays(rays_d)
        _need(counts_d, "counts_d", "u32", 2 * n)
        _check(lib().rtx_trace_counts(self._h, _dptr(rays_d), n, _dptr(counts_d),
                                      _stream_ptr(stream)), "rtx_trace_counts")


def debug_clip_slab(v9, region6, axis):
    """Host run of the builder's register-only triangle clip: (True, box6) when that path
    applies, (False, None) when the builder would run the full clip."""
    v = np.ascontiguousarray(v9, np.float32).reshape(9)
    r = np.ascontiguousarray(region6, np.float32).reshape(6)
    out = np.zeros(6, np.float32)
    rc = lib().rtx_debug_clip_slab(_p(v), _p(r), int(axis), _p(out))
    if rc < 0:
        _check(rc, "rtx_debug_clip_slab")
    return (True, out) if rc == 1 else (False, None)


def debug_recip_check(first=0, count=1 << 32):
    """The traversal's reciprocal against the IEEE quotient 1.0f / x on the GPU, over the
    float bit patterns [first, first + count): (mismatch count, lowest mismatching pattern
    or None)."""
    n = C.c_uint64(0)
    fb = C.c_uint32(0)
    _check(lib().rtx_debug_recip_check(int(first), int(count), C.byref(n), C.byref(fb)), "rtx_debug_recip_check")
    return n.value, (None if n.value == 0 else fb.value)


def debug_scratch_peak(reset=False):
    """Peak device scratch bytes held by library calls at once since the last reset."""
    v = C.c_uint64(0)
    _check(lib().rtx_debug_scratch_peak(C.byref(v), int(reset)), "rtx_debug_scratch_peak")
    return v.value


def build(verts6, tris):
    """GPU build from host arrays (verts (V,6) float32, tris (N,3) uint32)."""
    v = np.ascontiguousarray(verts6, np.float32).reshape(-1, 6)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    h = C.c_void_p()
    _check(lib().rtx_build(_p(v), len(v), _p(t), len(t), C.byref(h)), "rtx_build")
    return BVH(h)


def build_device(verts_d, tris_d, stream=None):
    """GPU build from device tensors (verts (V,6) float32, tris (N,3) int32/uint32)."""
    _need(verts_d, "verts_d")
    _need(tris_d, "tris_d", "u32")
    if verts_d.dim() != 2 or verts_d.shape[1] != 6 or tris_d.dim() != 2 or tris_d.shape[1] != 3:
        raise ValueError("verts_d must have shape (V, 6) and tris_d (N, 3)")
    h = C.c_void_p()
    _check(lib().rtx_build_device(_dptr(verts_d), verts_d.shape[0], _dptr(tris_d), tris_d.shape[0],
                                  _stream_ptr(stream), C.byref(h)), "rtx_build_device")
    return BVH(h)


def load_bvh(path):
    h = C.c_void_p()
    _check(lib().rtx_load(str(path).encode(), C.byref(h)), "rtx_load")
    return BVH(h)


def bake_bvh(scene_path):
    """Scene file -> GPU build -> <scene>.bvh, like bvh::bakeBVH."""
    _check(lib().rtx_bake_scene(str(scene_path).encode()), "rtx_bake_scene")


def load_scene(scene_path):
    v = C.POINTER(C.c_float)()
    t = C.POINTER(C.c_uint32)()
    nv, nt = C.c_uint32(), C.c_uint32()
    _check(lib().rtx_scene_load(str(scene_path).encode(), C.byref(v), C.byref(nv), C.byref(t), C.byref(nt)),
           "rtx_scene_load")
    verts = np.ctypeslib.as_array(v, shape=(nv.value, 6)).copy()
    tris = np.ctypeslib.as_array(t, shape=(nt.value, 3)).copy()
    lib().rtx_free_host(C.cast(v, C.c_void_p))
    lib().rtx_free_host(C.cast(t, C.c_void_p))
    return verts, tris


def geo_sphere(center, radius, rings=16, sectors=32):
    c = np.asarray(center, np.float32)
    nv, nt = C.c_uint32(), C.c_uint32()
    _check(lib().rtx_geo_sphere(_p(c), radius, rings, sectors, None, C.byref(nv), None, C.byref(nt)), "geo_sphere")
    v = np.zeros((nv.value, 6), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    _check(lib().rtx_geo_sphere(_p(c), radius, rings, sectors, _p(v), C.byref(nv), _p(t), C.byref(nt)), "geo_sphere")
    return v, t


def geo_box(center, half, quat_xyzw=(0, 0, 0, 1)):
    c = np.asarray(center, np.float32)
    h = np.asarray(half, np.float32)
    q = np.asarray(quat_xyzw, np.float32)
    nv, nt = C.c_uint32(), C.c_uint32()
    _check(lib().rtx_geo_box(_p(c), _p(h), _p(q), None, C.byref(nv), None, C.byref(nt)), "geo_box")
    v = np.zeros((nv.value, 6), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    _check(lib().rtx_geo_box(_p(c), _p(h), _p(q), _p(v), C.byref(nv), _p(t), C.byref(nt)), "geo_box")
    return v, t


def make_camera(cam):
    c = Camera()
    for k in ("pos", "fwd", "right", "up"):
        getattr(c, k)[:] = [float(x) for x in cam[k]]
    c.tan_half_fov_y = float(cam["tan_half_fov_y"])
    return c


def make_params(width, height, spp, max_bounce, rr_bounce=3, seed=0, env_mode=0, env_color=(0, 0, 0), ibl=None,
                uv=None, tex=None, sun=None):
    """ibl: (h, w, 3) float32 CUDA tensor, top row first, for env_mode 2. uv: (nverts, 2)
    and tex: (h, w, 3) float32 CUDA tensors for texture albedo. sun: (direction, cos of the
    cone, radiance) for env_mode 3. Tensors are kept alive on the returned params."""
    p = RenderParams()
    p.width, p.height, p.spp, p.max_bounce, p.rr_bounce, p.seed = width, height, spp, max_bounce, rr_bounce, seed
    p.env_mode = env_mode
    p.env_color[:] = [float(x) for x in env_color]
    p._keep = []
    if ibl is not None:
        _need(ibl, "ibl")
        if ibl.dim() != 3 or ibl.shape[2] != 3:
            raise ValueError("ibl must have shape (h, w, 3)")
        p._keep.append(ibl)
        p.ibl_d = ibl.data_ptr()
        p.ibl_height, p.ibl_width = int(ibl.shape[0]), int(ibl.shape[1])
    if tex is not None:
        set_texture(p, uv, tex)
    if sun is not None:
        d, c, rad = sun
        p.sun_dir[:] = [float(x) for x in d]
        p.sun_cos = float(c)
        p.sun_radiance[:] = [float(x) for x in rad]
    return p


def set_texture(p, uv, tex):
    """Point render params at texture albedo: uv (nverts, 2), tex (h, w, 3), contiguous
    float32 CUDA tensors (kept alive on the params)."""
    _need(uv, "uv")
    _need(tex, "tex")
    if uv.dim() != 2 or uv.shape[1] != 2 or tex.dim() != 3 or tex.shape[2] != 3:
        raise ValueError("uv must have shape (nverts, 2) and tex (h, w, 3)")
    p._tex_keep = (uv, tex)
    p.uv_d, p.tex_d = uv.data_ptr(), tex.data_ptr()
    p.tex_height, p.tex_width = int(tex.shape[0]), int(tex.shape[1])


def write_hdr(path, img, flip_y=True):
    """Radiance .hdr of an (h, w, 3) image; flip_y writes the bottom-up render top row first."""
    a = np.ascontiguousarray(img, np.float32)
    _check(lib().rtx_hdr_write(str(path).encode(), a.shape[1], a.shape[0], _p(a), int(flip_y)), "rtx_hdr_write")


def read_hdr(path):
    """(h, w, 3) float32 array of a Radiance .hdr file, top row first."""
    w, h = C.c_int(), C.c_int()
    ptr = C.POINTER(C.c_float)()
    _check(lib().rtx_hdr_read(str(path).encode(), C.byref(w), C.byref(h), C.byref(ptr)), "rtx_hdr_read")
    out = np.ctypeslib.as_array(ptr, shape=(h.value, w.value, 3)).copy()
    lib().rtx_free_host(C.cast(ptr, C.c_void_p))
    return out


# ---- path tracing ------------------------------------------------------------------

def get_tuning():
    t = Tuning()
    _check(lib().rtx_get_tuning(C.byref(t)), "rtx_get_tuning")
    return {k: getattr(t, k) for k, _ in Tuning._fields_}


class KernelTimer:
    """Times the traversal kernels of every trace call made inside the `with` block (HIP
    events on the caller's stream around the kernel alone, ray sorting excluded).
    After the block: .ms (summed kernel milliseconds) and .launches."""

    def __enter__(self):
        _check(lib().rtx_profile_begin(), "rtx_profile_begin")
        self.ms, self.launches = 0.0, 0
        return self

    def __exit__(self, *exc):
        ms, k = C.c_double(), C.c_uint32()
        _check(lib().rtx_profile_end(C.byref(ms), C.byref(k)), "rtx_profile_end")
        self.ms, self.launches = ms.value, k.value
        return False


def set_tuning(**kw):
    """Change launch shapes (see rtx_tuning in include/rtx.h); results never change."""
    cur = get_tuning()
    unknown = set(kw) - set(cur)
    if unknown:
        raise ValueError(f"unknown tuning keys {unknown}")
    cur.update({k: v for k, v in kw.items() if v is not None})
    t = Tuning(*[cur[k] for k, _ in Tuning._fields_])
    _check(lib().rtx_set_tuning(C.byref(t)), "rtx_set_tuning")


def _check_scene(bvh, params, albedo, emission):
    """Material and texture buffers against the tree they are indexed with."""
    inf = bvh.info()
    _need(albedo, "albedo", "f32", 3 * inf["input_tri_count"], optional=True)
    _need(emission, "emission", "f32", 3 * inf["input_tri_count"], optional=True)
    if params.tex_d:
        uv, tex = params._tex_keep
        if uv.shape[0] < inf["vert_count"]:
            raise ValueError(f"uv holds {uv.shape[0]} vertices, the tree has {inf['vert_count']}")


def render(bvh, cam, params, albedo=None, emission=None, out=None, stream=None):
    """Render into a (H, W, 3) float32 CUDA tensor. albedo/emission: (ntri, 3) CUDA tensors
    indexed by original triangle id (None = constant 0.8 / no emission)."""
    import torch
    _check_scene(bvh, params, albedo, emission)
    if out is None:
        out = torch.empty((params.height, params.width, 3), dtype=torch.float32, device="cuda")
    _need(out, "out", "f32", 3 * params.width * params.height)
    _check(lib().rtx_render(bvh.handle, C.byref(make_camera(cam)), C.byref(params), _dptr(albedo), _dptr(emission),
                            _dptr(out), _stream_ptr(stream)), "rtx_render")
    return out


def render_primal(bvh, cam, params, albedo=None, emission=None, out=None, stream=None):
    """render() that also returns each path's record ((2, spp * H * W, 4) float32 CUDA
    tensor, path g = sample * H * W + slot, the slot's pixel in 8 x 8 tile order (tile_pixel):
    [0, g] = its radiance and 0, [1, g] = its camera ray's hit t, u, v and triangle index
    bits): the primal that render_adjoint consumes."""
    import torch
    _check_scene(bvh, params, albedo, emission)
    if out is None:
        out = torch.empty((params.height, params.width, 3), dtype=torch.float32, device="cuda")
    _need(out, "out", "f32", 3 * params.width * params.height)
    paths = torch.empty((2, params.spp * params.height * params.width, 4), dtype=torch.float32, device="cuda")
    _check(lib().rtx_render_primal(bvh.handle, C.byref(make_camera(cam)), C.byref(params), _dptr(albedo),
                                   _dptr(emission), _dptr(out), _dptr(paths), _stream_ptr(stream)), "rtx_render_primal")
    return out, paths


def tile_pixel(k, width, height):
    """Pixel (y * width + x) of path slot k of a sample: 8 x 8 tiles, tile by tile along each
    band of 8 rows, row-major inside a tile (the renderer's path order; numpy arrays work)."""
    import numpy as np
    k = np.asarray(k, np.int64)
    band = k // (8 * width)
    y0 = 8 * band
    hb = np.minimum(8, height - y0)
    r = k - band * 8 * width
    full = width // 8
    in_full = r < full * 8 * hb
    t = r // (8 * hb)
    kk = r - t * 8 * hb
    wl = max(width - 8 * full, 1)
    k2 = r - full * 8 * hb
    row = np.where(in_full, kk // 8, k2 // wl)
    col = np.where(in_full, 8 * t + (kk & 7), 8 * full + k2 % wl)
    return (y0 + row) * width + col


def render_adjoint(bvh, cam, params, albedo, emission, d_image, grad_albedo=None, grad_emission=None, stream=None,
                   path_records=None):
    """Accumulate dLoss/dalbedo and dLoss/demission given dLoss/dimage (path replay). With
    texture albedo (params.tex_d set) grad_albedo is per texel and there is no emission
    gradient. path_records (from render_primal of the same render) saves the primal pass
    and the trace of the camera rays."""
    import torch
    textured = bool(params.tex_d)
    if path_records is not None:
        _need(path_records, "path_records", "f32", 8 * params.spp * params.height * params.width)
    _check_scene(bvh, params, albedo, emission)
    d_image = d_image.contiguous()
    _need(d_image, "d_image", "f32", 3 * params.width * params.height)
    if grad_albedo is None:
        grad_albedo = torch.zeros_like(params._tex_keep[1] if textured else albedo)
    if grad_emission is None and emission is not None and not textured:
        grad_emission = torch.zeros_like(emission)
    _need(grad_albedo, "grad_albedo", "f32",
          3 * (params.tex_width * params.tex_height if textured else bvh.info()["input_tri_count"]))
    _need(grad_emission, "grad_emission", "f32", 3 * bvh.info()["input_tri_count"], optional=True)
    if path_records is None:
        _check(lib().rtx_render_adjoint(bvh.handle, C.byref(make_camera(cam)), C.byref(params), _dptr(albedo),
                                        _dptr(emission), _dptr(d_image), _dptr(grad_albedo),
                                        _dptr(grad_emission), _stream_ptr(stream)), "rtx_render_adjoint")
    else:
        _check(lib().rtx_render_adjoint_replay(bvh.handle, C.byref(make_camera(cam)), C.byref(params), _dptr(albedo),
                                               _dptr(emission), _dptr(path_records), _dptr(d_image),
                                               _dptr(grad_albedo), _dptr(grad_emission), _stream_ptr(stream)),
               "rtx_render_adjoint_replay")
    return grad_albedo, grad_emission


class Batch:
    """Several scenes rendered together; mode 0 plain (one launch sequence per scene),
    1 batched and captured once in a hipGraph, 2 batched without the graph."""

    def __init__(self, bvhs, cams, params, albedos, emissions, images, mode):
        n = len(bvhs)
        if not (len(cams) == len(params) == len(albedos) == len(emissions) == len(images) == n):
            raise ValueError("one camera, params, albedo, emission and image per scene")
        for b, p, a, e, im in zip(bvhs, params, albedos, emissions, images):
            _check_scene(b, p, a, e)
            _need(im, "image", "f32", 3 * p.width * p.height)
        self._keep = (bvhs, albedos, emissions, images, params)   # params hold IBL/texture tensors
        hs = (C.c_void_p * n)(*[b.handle for b in bvhs])
        cs = (Camera * n)(*[make_camera(c) for c in cams])
        ps = (RenderParams * n)(*params)
        al = (C.c_void_p * n)(*[None if a is None else a.data_ptr() for a in albedos])
        em = (C.c_void_p * n)(*[None if e is None else e.data_ptr() for e in emissions])
        im = (C.c_void_p * n)(*[i.data_ptr() for i in images])
        h = C.c_void_p()
        _check(lib().rtx_batch_create(hs, n, cs, ps, al, em, im, mode, C.byref(h)), "rtx_batch_create")
        self._h = h

    def launch(self, stream=None):
        _check(lib().rtx_batch_launch(self._h, _stream_ptr(stream)), "rtx_batch_launch")

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.rtx_batch_free(self._h)
            self._h = None


def _autograd_fn():
    import torch

    class RenderFunction(torch.autograd.Function):
        """image = render(albedo, emission); the forward keeps each path's record (radiance
        and camera-ray hit, 32 bytes per path, whatever max_bounce) and backward re-traces the
        paths from their first hits (path replay)."""

        @staticmethod
        def forward(ctx, albedo, emission, bvh, cam, params):
            img, paths = render_primal(bvh, cam, params, albedo, emission)
            ctx.save_for_backward(albedo, emission, paths)
            ctx.scene = (bvh, cam, params)
            return img

        @staticmethod
        def backward(ctx, grad_img):
            albedo, emission, paths = ctx.saved_tensors
            bvh, cam, params = ctx.scene
            ga, ge = render_adjoint(bvh, cam, params, albedo, emission, grad_img.contiguous(), path_records=paths)
            return ga, ge, None, None, None

    return RenderFunction


def _texture_autograd_fn():
    import torch

    class TextureRenderFunction(torch.autograd.Function):
        """image = render with albedo from a texture; backward re-traces the paths and
        accumulates the gradient per texel."""

        @staticmethod
        def forward(ctx, tex, uv, emission, bvh, cam, params):
            tex_c = tex.detach().contiguous()
            set_texture(params, uv, tex_c)
            img, paths = render_primal(bvh, cam, params, None, emission)
            ctx.save_for_backward(tex_c, paths)
            ctx.scene = (uv, emission, bvh, cam, params)
            return img

        @staticmethod
        def backward(ctx, grad_img):
            tex_c, paths = ctx.saved_tensors
            uv, emission, bvh, cam, params = ctx.scene
            set_texture(params, uv, tex_c)
            gtex, _ = render_adjoint(bvh, cam, params, None, emission, grad_img.contiguous(), path_records=paths)
            return gtex, None, None, None, None, None

    return TextureRenderFunction


_RF = None
_TF = None


def differentiable_render_texture(tex, uv, bvh, cam, params, emission=None):
    """torch.autograd entry for texture albedo: gradients flow to tex ((h, w, 3) CUDA
    tensor; uv (nverts, 2)). Requires params.rr_bounce > params.max_bounce. Gradients are
    exact for texels > 0; a zero texel channel gets no gradient from the path vertices on
    it, so keep the fitted texture in (0, 1]."""
    global _TF
    if _TF is None:
        _TF = _texture_autograd_fn()
    return _TF.apply(tex, uv, emission, bvh, cam, params)


def differentiable_render(albedo, emission, bvh, cam, params):
    """torch.autograd entry: gradients flow to albedo and emission ((ntri, 3) CUDA tensors).
    Requires params.rr_bounce > params.max_bounce. Albedo gradients are exact for albedo > 0
    (a zero albedo channel gets none from the vertices on it)."""
    global _RF
    if _RF is None:
        _RF = _autograd_fn()
    return _RF.apply(albedo, emission, bvh, cam, params)
