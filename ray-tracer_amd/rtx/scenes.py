"""Synthetic workloads for the hot path (host side, numpy).

Vertices use the reference Vertex layout (src/bvh/bvh.hpp:32-41): 6 float32 per
vertex (x, y, z, nx, ny, nz). Triangles are 3 uint32 vertex indices. Rays are 8
float32: (ox, oy, oz, tmin, dx, dy, dz, tmax).

Every generator is deterministic for a given seed; the GPU path and the CPU oracle
always consume the same arrays, produced once per run.
"""
import math

import numpy as np

F32 = np.float32


def _normalize(v):
    v = np.asarray(v, F32)
    n = np.atleast_1d(np.sqrt((v * v).sum(-1, dtype=F32)).astype(F32))
    n[n == 0] = 1
    return (v / n.reshape(v.shape[:-1] + (1,))).astype(F32)


def random_triangles(n, seed=1, mu=-5.0, sigma=0.5):
    """Triangle soup: centres uniform in the unit cube, Gaussian edge vectors with a
    lognormal(mu, sigma) scale, unshared vertices, face normals as vertex normals."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3))
    s = np.exp(rng.normal(mu, sigma, n))[:, None]
    e1 = rng.standard_normal((n, 3)) * s
    e2 = rng.standard_normal((n, 3)) * s
    v0 = c - (e1 + e2) / 3.0
    v1 = v0 + e1
    v2 = v0 + e2
    pos = np.stack([v0, v1, v2], 1).astype(F32)            # (n, 3, 3)
    fn = _normalize(np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0]))
    verts = np.empty((n, 3, 6), F32)
    verts[:, :, :3] = pos
    verts[:, :, 3:] = fn[:, None, :]
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return verts.reshape(-1, 6), tris


def incoherent_rays(n, seed=1, lo=0.0, hi=1.0):
    """Origins uniform in [lo,hi]^3, directions uniform on the sphere, tmin 0, tmax inf."""
    rng = np.random.Generator(np.random.PCG64(seed + 0x5bd1e995))
    o = lo + (hi - lo) * rng.random((n, 3))
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.zeros((n, 8), F32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    return rays


def camera_lookat(pos, target, fov_deg):
    """Pinhole camera of makeCameraLookAt (src/main.cpp:55-66) in float32."""
    pos = np.asarray(pos, F32)
    target = np.asarray(target, F32)
    fwd = _normalize(target - pos)
    world_up = np.array([0, 1, 0], F32) if abs(fwd[1]) < F32(0.99) else np.array([1, 0, 0], F32)
    right = _normalize(np.cross(fwd, world_up))
    up = np.cross(right, fwd).astype(F32)
    tan_h = F32(math.tan(F32(fov_deg) * F32(math.pi) / F32(180.0) * F32(0.5)))
    return dict(pos=pos, fwd=fwd, right=right, up=up, tan_half_fov_y=tan_h)


def pinhole_rays(cam, width, height, jitter=0.5, tmin=1e-3):
    """One primary ray per pixel through the pixel position (x + jitter, y + jitter),
    following primaryRay (src/main.cpp:69-83)."""
    ys, xs = np.meshgrid(np.arange(height, dtype=F32), np.arange(width, dtype=F32), indexing="ij")
    aspect = F32(width) / F32(height)
    u = ((xs + F32(jitter)) / F32(width) * F32(2) - F32(1)) * aspect
    v = (ys + F32(jitter)) / F32(height) * F32(2) - F32(1)
    th = F32(cam["tan_half_fov_y"])
    d = (cam["fwd"][None, None, :] + cam["right"][None, None, :] * (u * th)[..., None]
         + cam["up"][None, None, :] * (v * th)[..., None]).astype(F32)
    d = _normalize(d.reshape(-1, 3))
    rays = np.zeros((width * height, 8), F32)
    rays[:, 0:3] = cam["pos"]
    rays[:, 3] = tmin
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    return rays


def icosphere(subdiv, radius=1.0, center=(0, 0, 0)):
    """Geodesic sphere: icosahedron with `subdiv` midpoint subdivisions (20*4^s faces),
    visible from outside under the reference's back-face test, with radial normals."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.array(verts)
    out = np.zeros((len(P), 6), F32)
    out[:, :3] = P * radius + np.asarray(center)
    out[:, 3:] = P
    tris = np.array(faces, np.int64)
    # Visible from outside: e1 x e2 must point inwards (see quad()).
    e1 = P[tris[:, 1]] - P[tris[:, 0]]
    e2 = P[tris[:, 2]] - P[tris[:, 0]]
    outward = (np.cross(e1, e2) * P[tris].mean(1)).sum(-1) > 0
    tris[outward] = tris[outward][:, [0, 2, 1]]
    return out, tris.astype(np.uint32)


_ICO_T = (1.0 + 5 ** 0.5) / 2.0
_ICO_V = [(-1, _ICO_T, 0), (1, _ICO_T, 0), (-1, -_ICO_T, 0), (1, -_ICO_T, 0), (0, -1, _ICO_T), (0, 1, _ICO_T),
          (0, -1, -_ICO_T), (0, 1, -_ICO_T), (_ICO_T, 0, -1), (_ICO_T, 0, 1), (-_ICO_T, 0, -1), (-_ICO_T, 0, 1)]
_ICO_F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
          (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
          (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]


def geodesic_sphere(freq, radius=1.0, center=(0, 0, 0), displace=0.0, seed=0):
    """Class-I geodesic icosphere of frequency `freq`: every icosahedron face is cut into a
    freq x freq triangular grid (20 freq^2 triangles, 10 freq^2 + 2 shared vertices) projected
    onto the sphere, optionally displaced along the radius by displace * fBm. Shading normals
    are the undisplaced radial ones; triangles face outwards under the reference's back-face
    test (see quad())."""
    V = np.array(_ICO_V, np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(_ICO_F, np.int64)
    n = int(freq)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    m = (ii + jj) <= n
    i, j = ii[m], jj[m]                               # grid points of one face, weights (n-i-j, i, j)
    local = -np.ones((n + 1, n + 1), np.int64)
    local[i, j] = np.arange(len(i))
    # A point is the integer combination sum_k w_k V_k over the face's corners; the same
    # point on a shared edge or corner has the same (vertex, weight) pairs from every face.
    w = np.stack([n - i - j, i, j], 1)                # (P, 3)
    keys, pos = [], []
    for a, b, c in F:
        vid = np.array([a, b, c])[None, :].repeat(len(i), 0)
        vid = np.where(w > 0, vid, 99)                # corners with weight 0 do not count
        order = np.argsort(vid, 1, kind="stable")
        keys.append(np.concatenate([np.take_along_axis(vid, order, 1), np.take_along_axis(w, order, 1)], 1))
        pos.append((w[:, 0:1] * V[a] + w[:, 1:2] * V[b] + w[:, 2:3] * V[c]) / n)
    keys = np.concatenate(keys)
    pos = np.concatenate(pos)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    P = pos[first]
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    # triangles of one face grid: up (i,j)(i+1,j)(i,j+1) and down (i+1,j)(i+1,j+1)(i,j+1)
    up = (i + j) <= n - 1
    dn = (i + j) <= n - 2
    t_up = np.stack([local[i[up], j[up]], local[i[up] + 1, j[up]], local[i[up], j[up] + 1]], 1)
    t_dn = np.stack([local[i[dn] + 1, j[dn]], local[i[dn] + 1, j[dn] + 1], local[i[dn], j[dn] + 1]], 1)
    tl = np.concatenate([t_up, t_dn])
    tris = np.concatenate([inv[tl + f * len(i)] for f in range(len(F))])
    r = radius * (1.0 + displace * fbm(P, seed=seed)) if displace else np.full(len(P), radius)
    verts = np.zeros((len(P), 6), F32)
    verts[:, :3] = P * r[:, None] + np.asarray(center)
    verts[:, 3:] = P
    e1 = P[tris[:, 1]] - P[tris[:, 0]]
    e2 = P[tris[:, 2]] - P[tris[:, 0]]
    outward = (np.cross(e1, e2) * P[tris].mean(1)).sum(-1) > 0
    tris[outward] = tris[outward][:, [0, 2, 1]]
    return verts, tris.astype(np.uint32)


def merge(parts):
    """Concatenate (verts6, tris) meshes, offsetting indices."""
    vs, ts, off = [], [], 0
    for v, t in parts:
        vs.append(np.asarray(v, F32))
        ts.append(np.asarray(t, np.uint32) + off)
        off += len(v)
    return np.concatenate(vs), np.concatenate(ts)


def quad(p0, p1, p2, p3, facing):
    """Two triangles (p0,p1,p2), (p0,p2,p3) seen from the side `facing` points to.

    The reference culls a = dot(e1, cross(d, e2)) >= 0 (src/bvh/bvh_traverse.cpp:57-60),
    so a face is hit only by rays travelling along e1 x e2. The winding is chosen so
    that e1 x e2 points away from the visible side; the shading normal is `facing`."""
    P = np.array([p0, p1, p2, p3], F32)
    f = np.asarray(facing, F32)
    if float(np.dot(np.cross(P[1] - P[0], P[2] - P[0]), f)) > 0:
        P = P[[0, 3, 2, 1]]
    v = np.zeros((4, 6), F32)
    v[:, :3] = P
    v[:, 3:] = _normalize(f[None])[0]
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def cornell_box(seed=0, sphere_subdiv=3, sphere_radius=0.3, light_radiance=15.0, albedo_lo=0.2, albedo_hi=0.8,
                sphere_center=(0.0, 0.0, 0.0)):
    """Open-front Cornell-style box in [-1,1]^3: 10 wall quads (20 triangles, inward
    facing), one ceiling light quad (2 triangles) and an icosphere, centred by default.
    Returns verts6, tris, albedo (3 per triangle), emission (3 per triangle), camera."""
    parts = []
    # floor y=-1, ceiling y=1, back z=1, left x=-1, right x=1, each as 2 quads (split in x or z/y)
    def wall(fn, facing):
        for i in range(2):
            parts.append(quad(*fn(i), facing))
    wall(lambda i: [(-1 + i, -1, -1), (-1 + i, -1, 1), (i, -1, 1), (i, -1, -1)], (0, 1, 0))    # floor
    wall(lambda i: [(-1 + i, 1, -1), (i, 1, -1), (i, 1, 1), (-1 + i, 1, 1)], (0, -1, 0))      # ceiling
    wall(lambda i: [(-1 + i, -1, 1), (-1 + i, 1, 1), (i, 1, 1), (i, -1, 1)], (0, 0, -1))      # back
    wall(lambda i: [(-1, -1, -1 + i), (-1, 1, -1 + i), (-1, 1, i), (-1, -1, i)], (1, 0, 0))   # left
    wall(lambda i: [(1, -1, -1 + i), (1, -1, i), (1, 1, i), (1, 1, -1 + i)], (-1, 0, 0))      # right
    # ceiling light slightly below the ceiling, facing down
    parts.append(quad((-0.25, 0.99, -0.25), (0.25, 0.99, -0.25), (0.25, 0.99, 0.25), (-0.25, 0.99, 0.25),
                      (0, -1, 0)))
    sv, st = icosphere(sphere_subdiv, sphere_radius, sphere_center)
    parts.append((sv, st))
    verts, tris = merge(parts)
    nt = len(tris)
    rng = np.random.Generator(np.random.PCG64(seed))
    albedo = rng.uniform(albedo_lo, albedo_hi, (nt, 3)).astype(F32)
    emission = np.zeros((nt, 3), F32)
    emission[20:22] = light_radiance
    cam = camera_lookat((0, 0, -3.2), (0, 0, 0), 40.0)
    return verts, tris, albedo, emission, cam


def fbm(p, octaves=4, seed=0):
    """Smooth deterministic fBm on 3-D points: a sum of octaves of randomly oriented sine
    waves with halving amplitude and doubling frequency. Returns values in about [-1, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed + 77))
    out = np.zeros(len(p))
    amp, freq, norm = 1.0, 2.0, 0.0
    for _ in range(octaves):
        for _ in range(3):
            k = rng.standard_normal(3)
            k /= np.linalg.norm(k)
            out += amp * np.sin(freq * (p @ k) + rng.uniform(0, 2 * math.pi)) / 3.0
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return out / norm


def uv_sphere(rings, sectors, radius=1.0, center=(0, 0, 0), displace=0.0, seed=0):
    """UV sphere with the reference's vertex/triangle layout (src/bvh/bvh_geo.cpp:22-112:
    2 + (rings-1)*sectors vertices, 2*sectors*(rings-1) triangles), optionally displaced
    along the normal by displace * fBm. Shading normals are the undisplaced radial ones."""
    th = np.pi * np.arange(1, rings) / rings
    ph = 2 * np.pi * np.arange(sectors) / sectors
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(Pp), np.cos(T), np.sin(T) * np.sin(Pp)], -1).reshape(-1, 3)
    n = np.concatenate([[[0, 1, 0]], n, [[0, -1, 0]]])
    r = radius * (1.0 + displace * fbm(n, seed=seed)) if displace else np.full(len(n), radius)
    verts = np.zeros((len(n), 6), F32)
    verts[:, :3] = n * r[:, None] + np.asarray(center)
    verts[:, 3:] = n
    south = len(n) - 1

    def ring(rr, s):
        return 1 + (rr - 1) * sectors + (s % sectors)
    tris = []
    for s in range(sectors):
        tris.append((0, ring(1, s + 1), ring(1, s)))
    for rr in range(1, rings - 1):
        for s in range(sectors):
            a, b, c, d = ring(rr, s), ring(rr, s + 1), ring(rr + 1, s + 1), ring(rr + 1, s)
            tris += [(a, b, c), (a, c, d)]
    for s in range(sectors):
        tris.append((south, ring(rings - 1, s + 1), ring(rings - 1, s)))
    tris = np.array(tris, np.int64)
    # visible from outside under the reference's culling rule (see quad())
    P = verts[:, :3].astype(np.float64)
    cr = np.cross(P[tris[:, 1]] - P[tris[:, 0]], P[tris[:, 2]] - P[tris[:, 0]])
    out = (cr * (P[tris].mean(1) - np.asarray(center))).sum(-1) > 0
    tris[out] = tris[out][:, [0, 2, 1]]
    return verts, tris.astype(np.uint32)


def diff_scene(seed=3, rings=33, sectors=64, displace=0.08, albedo_lo=0.1, albedo_hi=0.9):
    """Differentiable-render scene: the Cornell box walls, a displaced UV sphere of
    2*sectors*(rings-1) triangles (4096 for 33 x 64) and a 2-triangle ceiling emitter.
    Returns verts, tris, albedo (per triangle, seeded), emission, camera."""
    parts = []

    def wall(fn, facing):
        for i in range(2):
            parts.append(quad(*fn(i), facing))
    wall(lambda i: [(-1 + i, -1, -1), (-1 + i, -1, 1), (i, -1, 1), (i, -1, -1)], (0, 1, 0))
    wall(lambda i: [(-1 + i, 1, -1), (i, 1, -1), (i, 1, 1), (-1 + i, 1, 1)], (0, -1, 0))
    wall(lambda i: [(-1 + i, -1, 1), (-1 + i, 1, 1), (i, 1, 1), (i, -1, 1)], (0, 0, -1))
    wall(lambda i: [(-1, -1, -1 + i), (-1, 1, -1 + i), (-1, 1, i), (-1, -1, i)], (1, 0, 0))
    wall(lambda i: [(1, -1, -1 + i), (1, -1, i), (1, 1, i), (1, 1, -1 + i)], (-1, 0, 0))
    parts.append(quad((-0.3, 0.98, -0.3), (0.3, 0.98, -0.3), (0.3, 0.98, 0.3), (-0.3, 0.98, 0.3), (0, -1, 0)))
    parts.append(uv_sphere(rings, sectors, 0.45, (0, -0.3, 0.2), displace, seed))
    verts, tris = merge(parts)
    rng = np.random.Generator(np.random.PCG64(seed))
    albedo = rng.uniform(albedo_lo, albedo_hi, (len(tris), 3)).astype(F32)
    emission = np.zeros((len(tris), 3), F32)
    emission[20:22] = 12.0
    cam = camera_lookat((0, 0, -3.0), (0, -0.1, 0), 42.0)
    return verts, tris, albedo, emission, cam


def batched_scenes(n=64, seed=2, lo=1000, hi=500_000, displace=0.05):
    """Scenes of uneven size: fBm-displaced geodesic icospheres whose triangle counts
    20 freq^2 follow a lognormal clamped to [lo, hi] (freq 8..158: 1,280 to 499,280
    triangles); resolutions from {128, 256, 512}^2; spp uniform in 8..64; max_bounce 6.
    Returns a list of dicts (verts, tris, cam, width, height, spp, max_bounce)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    fmin, fmax = math.ceil(math.sqrt(lo / 20.0)), math.floor(math.sqrt(hi / 20.0))
    out = []
    for i in range(n):
        ntri = int(np.clip(np.exp(rng.normal(math.log(20000), 1.2)), lo, hi))
        freq = int(np.clip(round(math.sqrt(ntri / 20.0)), fmin, fmax))
        res = int(rng.choice([128, 256, 512]))
        spp = int(rng.integers(8, 65))
        v, t = geodesic_sphere(freq, 1.0, (0, 0, 0), displace, seed=seed * 1000 + i)
        cam = camera_lookat((0.3 * math.sin(i), 0.4, -3.2), (0, 0, 0), 45.0)
        out.append(dict(verts=v, tris=t, cam=cam, width=res, height=res, spp=spp, max_bounce=6))
    return out


def _value_noise(rng, res, x, z):
    """Bilinear value noise on a res x res lattice over [0,1]^2 with smoothstep weights."""
    g = rng.random((res + 1, res + 1))
    fx, fz = x * res, z * res
    ix = np.minimum(fx.astype(np.int64), res - 1)
    iz = np.minimum(fz.astype(np.int64), res - 1)
    tx, tz = fx - ix, fz - iz
    sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
    a = g[ix, iz] * (1 - sx) + g[ix + 1, iz] * sx
    b = g[ix, iz + 1] * (1 - sx) + g[ix + 1, iz + 1] * sx
    return a * (1 - sz) + b * sz


def fbm2(n, octaves=6, seed=4, base_res=4):
    """n x n fBm height samples over [0,1]^2 (lacunarity 2, gain 0.5), values in about [0,1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x, z = np.meshgrid(np.linspace(0.0, 1.0, n), np.linspace(0.0, 1.0, n), indexing="ij")
    h = np.zeros((n, n))
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        h += amp * _value_noise(rng, base_res << o, x, z)
        tot += amp
        amp *= 0.5
    return h / tot


def terrain(n=2048, octaves=6, seed=4, height=0.15):
    """Heightfield terrain over [0,1] x [0,1] in x/z (y up): n x n vertices from fBm, two
    triangles per cell (2 (n-1)^2 triangles) wound for the reference's back-face test so
    they are seen from above, vertex normals from central differences, uv = (x, z).
    Returns verts6 (n*n, 6), tris (2 (n-1)^2, 3), uv (n*n, 2)."""
    h = fbm2(n, octaves, seed) * height
    xs = np.linspace(0.0, 1.0, n)
    x, z = np.meshgrid(xs, xs, indexing="ij")
    dx = np.gradient(h, xs, axis=0)
    dz = np.gradient(h, xs, axis=1)
    nrm = np.stack([-dx, np.ones_like(h), -dz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    verts = np.empty((n, n, 6), F32)
    verts[..., 0], verts[..., 1], verts[..., 2] = x, h, z
    verts[..., 3:] = nrm
    idx = np.arange(n * n, dtype=np.uint32).reshape(n, n)
    a, b = idx[:-1, :-1], idx[1:, :-1]          # (i, j), (i+1, j)
    c, d = idx[:-1, 1:], idx[1:, 1:]            # (i, j+1), (i+1, j+1)
    t1 = np.stack([a, b, c], -1).reshape(-1, 3)
    t2 = np.stack([b, d, c], -1).reshape(-1, 3)
    tris = np.concatenate([t1, t2], 0).astype(np.uint32)
    uv = np.stack([x, z], -1).reshape(-1, 2).astype(F32)
    return verts.reshape(-1, 6), tris, uv


def albedo_texture(res=512, seed=4):
    """Ground-truth RGB albedo texture (res x res x 3, values in [0.1, 0.9]): per-channel
    fBm patterns."""
    ch = [fbm2(res, 4, seed * 7 + k, base_res=3) for k in range(3)]
    t = np.stack(ch, -1)
    t = (t - t.min()) / max(t.max() - t.min(), 1e-6)
    return (0.1 + 0.8 * t).astype(F32)


def terrain_setup(n=2048, res=(1920, 1080), seed=4):
    """Camera over the terrain and a sun + sky environment (env_mode 3):
    (camera, sky colour, (sun direction, cos of the cone, sun radiance))."""
    cam = camera_lookat((0.5, 0.7, 0.15), (0.5, 0.0, 0.55), 35.0)
    sun_dir = _normalize(np.array([[0.35, 1.0, 0.25]]))[0]
    sun = (tuple(float(x) for x in sun_dir), float(np.cos(np.radians(6.0))), (25.0, 22.0, 18.0))
    return cam, (0.35, 0.45, 0.65), sun
