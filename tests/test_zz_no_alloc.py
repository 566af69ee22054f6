"""rtx_trace_closest_noalloc (no allocation: scratch in a fixed device array) against
rtx_trace_closest, through the C ABI, on two calls in a row (the second smaller), with the
ray sort and the persistent kernel both in use; and its size limit."""
import ctypes as C

import numpy as np
import pytest

import rtx
from rtx import scenes as S

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 21


def _run(fn, bvh, rays_d, n, hits_d, nrm_d):
    return fn(bvh.handle, C.c_void_p(rays_d.data_ptr()), n, C.c_void_p(hits_d.data_ptr()),
              None if nrm_d is None else C.c_void_p(nrm_d.data_ptr()), None)


def test_noalloc_matches_trace_closest():
    import torch
    L = rtx.lib()
    v, t = S.random_triangles(200_000, seed=5)
    g = rtx.build(v, t)
    saved = rtx.get_tuning()
    try:
        rtx.set_tuning(trace_loop=2, ray_sort=1, leaf_coop=1)
        for n, normals in ((300_000, True), (70_001, False), (5_000, True)):
            print(f"[noalloc] {n} rays, normals={normals}", flush=True)
            rays_d = torch.from_numpy(S.incoherent_rays(n, seed=n)).cuda()
            ref = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
            out = torch.full((n, 4), -7.0, dtype=torch.float32, device="cuda")
            rn = torch.zeros((n, 3), dtype=torch.float32, device="cuda") if normals else None
            on = torch.full((n, 3), -7.0, dtype=torch.float32, device="cuda") if normals else None
            assert _run(L.rtx_trace_closest, g, rays_d, n, ref, rn) == 0
            assert _run(L.rtx_trace_closest_noalloc, g, rays_d, n, out, on) == 0
            torch.cuda.synchronize()
            a, b = ref.cpu().numpy(), out.cpu().numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), n
            assert (a[:, 3].view(np.uint32) != 0xFFFFFFFF).any()
            if normals:
                assert np.array_equal(rn.cpu().numpy().view(np.uint32), on.cpu().numpy().view(np.uint32))
    finally:
        rtx.set_tuning(**saved)


def test_noalloc_limit_and_empty():
    import torch
    L = rtx.lib()
    v, t = S.random_triangles(1000, seed=2)
    g = rtx.build(v, t)
    n = MAX_RAYS + 1
    rays_d = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    hits_d = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    assert _run(L.rtx_trace_closest_noalloc, g, rays_d, n, hits_d, None) == -5   # RTX_E_LIMIT, nothing launched
    assert b"2097152" in L.rtx_last_error()
    assert L.rtx_trace_closest_noalloc(g.handle, None, 0, None, None, None) == 0
