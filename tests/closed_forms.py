"""Closed-form checks of the path-trace loop's random walk (src/main.cpp:123-141, 145-204).

Each check renders a scene whose expected image statistics follow from the mathematics
of the loop, not from its code, through a caller-supplied `render` function (the CPU
oracle or the GPU renderer), and reports the measured value, the expectation and the
statistical tolerance:

* sky_plane: a one-sided plane of albedo a under the reference's sky gradient, one
  bounce. The bounce direction d follows the cosine lobe around the plane normal n, whose
  moments are E[d] = (2/3) n, E[(d.t)^2] = E[(d.b)^2] = 1/4, E[(d.n)^2] = 1/2 for tangents
  t, b. The sky is linear in d.y, so each channel's image mean is
  a (1 + (k - 1)(E[d.y]/2 + 1/2)) with k the zenith colour (0.5, 0.7, 1).
* ibl_bands: the same plane under an equirectangular map of indicator bands (rows = polar
  angle, columns = azimuth). Each band's expected energy is the cosine lobe's integral
  over the band, computed by quadrature on the sphere (for n = +y: the sin^2 differences
  of the rows, uniform columns).
* roulette_furnace: a closed box, emission 1 and albedo (0.3, 0.6, 0.9) on every wall,
  roulette from bounce 0. With survival probability p = max(throughput) the blue
  throughput is renormalized to 1 and survives each bounce with probability 0.9, so the
  number of survived bounces S is geometric, truncated at max_bounce. The image mean is
  E[L_c] = sum_k a_c^k (roulette is unbiased), and every moment of L_c follows from the
  distribution of S.

Statistics use one path per pixel sample; tolerances are `nsig` standard errors.
"""
import math

import numpy as np

from rtx import scenes as S

ZENITH = np.array([0.5, 0.7, 1.0])


class Result:
    def __init__(self, name, got, exp, tol):
        self.name = name
        self.got = np.atleast_1d(np.asarray(got, np.float64))
        self.exp = np.atleast_1d(np.asarray(exp, np.float64))
        self.tol = np.atleast_1d(np.asarray(tol, np.float64))
        self.ok = bool(np.all(np.abs(self.got - self.exp) <= self.tol))

    def __repr__(self):
        dev = np.abs(self.got - self.exp) / np.maximum(self.tol, 1e-300)
        return (f"{self.name}: got {np.array2string(self.got, precision=5)} expected "
                f"{np.array2string(self.exp, precision=5)} (worst {dev.max():.2f} of the tolerance)")


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def plane(normal, half=100.0):
    """One quad through the origin seen from the side `normal` points to. Its centre is
    offset so that the diagonal shared by its two triangles stays far from the origin:
    Moller-Trumbore is not watertight, and a ray can pass between two triangles there."""
    n = _unit(normal)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.5 else np.array([0.0, 0.0, 1.0])
    t1 = _unit(np.cross(n, a))
    t2 = np.cross(n, t1)
    c = 0.5 * half * (t1 - t2)
    pts = [c + half * (-t1 - t2), c + half * (t1 - t2), c + half * (t1 + t2), c + half * (-t1 + t2)]
    return S.quad(*pts, n.astype(np.float32))


def plane_camera(normal):
    """Camera one unit above the plane's centre, looking at it with a 20-degree view."""
    n = _unit(normal)
    return S.camera_lookat(tuple(n), (0.0, 0.0, 0.0), 20.0)


def shading_normal(normal):
    """The normal the renderer shades with: the float32 vertex normal, normalized."""
    n = np.asarray(normal, np.float32)
    return (n / np.sqrt((n * n).sum(dtype=np.float32))).astype(np.float64)


def sky_plane(render, normal, a=0.7, W=64, H=64, spp=16, nsig=5.0):
    """Image mean of the plane under the sky gradient with max_bounce 1 (see module doc).
    Returns (colour result, blue-exactness result)."""
    v, t = plane(normal)
    n = shading_normal(normal)
    albedo = np.full((len(t), 3), a, np.float32)
    img = render(v, t, albedo, None, plane_camera(normal), W, H, spp, 1, 99, 11)
    N = W * H * spp
    got = img.reshape(-1, 3).astype(np.float64).mean(0)
    ey = 2.0 / 3.0 * n[1]
    var_y = (1.0 - n[1] ** 2) / 4.0 + n[1] ** 2 / 2.0 - ey ** 2
    exp = a * (1.0 + (ZENITH - 1.0) * (ey / 2.0 + 0.5))
    sig = a * np.abs(ZENITH - 1.0) / 2.0 * math.sqrt(var_y / N)
    r = Result(f"sky plane n={np.round(n, 3).tolist()}", got[:2], exp[:2], nsig * sig[:2])
    # blue: the sky's blue is 1 in every direction, so every pixel is exactly a
    b = Result("sky plane blue", img[..., 2].astype(np.float64).max(), a, 1e-6 * a)
    b.ok = b.ok and bool(np.all(np.abs(img[..., 2].astype(np.float64) - a) <= 1e-6 * a))
    return r, b


def lobe_band_energy(n, h, w, m=160):
    """Cosine-lobe probability of every texel of an h x w equirectangular map (row 0 =
    straight up, column 0 starts at azimuth -pi; IBL::sample, main.cpp:107-117), by the
    midpoint rule on m x m points per texel of the integrand max(0, d.n)/pi sin(theta)."""
    n = _unit(n)
    th = (np.arange(h * m) + 0.5) * (math.pi / (h * m))
    ph = -math.pi + (np.arange(w * m) + 0.5) * (2.0 * math.pi / (w * m))
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    cosn = st * np.cos(ph)[None, :] * n[0] + ct * n[1] + st * np.sin(ph)[None, :] * n[2]
    f = np.maximum(cosn, 0.0) / math.pi * st * (math.pi / (h * m)) * (2.0 * math.pi / (w * m))
    return f.reshape(h, m, w, m).sum((1, 3))


def ibl_bands(render, normal, a=0.7, W=64, H=64, spp=16, h=8, w=16, nsig=5.0):
    """Energy per row band and per column band of an indicator map (see module doc): three
    bands per render, one per colour channel. Returns one Result over all bands."""
    v, t = plane(normal)
    n = shading_normal(normal)
    albedo = np.full((len(t), 3), a, np.float32)
    cam = plane_camera(normal)
    E = lobe_band_energy(n, h, w)
    masks, probs = [], []
    for r in range(h):
        m = np.zeros((h, w), bool); m[r, :] = True
        masks.append(m); probs.append(E[r].sum())
    for c in range(w):
        m = np.zeros((h, w), bool); m[:, c] = True
        masks.append(m); probs.append(E[:, c].sum())
    N = W * H * spp
    got = []
    for k in range(0, len(masks), 3):
        ibl = np.zeros((h, w, 3), np.float32)
        for ch, mk in enumerate(masks[k:k + 3]):
            ibl[mk, ch] = 1.0
        img = render(v, t, albedo, None, cam, W, H, spp, 1, 99, 13, env_mode=2, ibl=ibl)
        mean = img.reshape(-1, 3).astype(np.float64).mean(0)
        got += list(mean[:len(masks[k:k + 3])])
    p = np.array(probs)
    sig = a * np.sqrt(p * (1 - p) / N)
    return Result(f"IBL bands n={np.round(n, 3).tolist()}", got, a * p, nsig * sig + 1e-9)


FURNACE_ALBEDO = np.array([0.3, 0.6, 0.9])


def closed_box():
    """Six inward-facing quads bounding [-1,1]^3: every path hits a wall at every bounce."""
    parts = []
    for axis in range(3):
        for sgn in (-1, 1):
            c = np.zeros(3)
            c[axis] = sgn
            u = np.zeros(3); u[(axis + 1) % 3] = 1
            w = np.zeros(3); w[(axis + 2) % 3] = 1
            parts.append(S.quad(c - u - w, c + u - w, c + u + w, c - u + w, -c))
    return S.merge(parts)


def furnace_moments(max_bounce, a=FURNACE_ALBEDO, keep=0.9):
    """Mean and variance of each channel of L = sum_{k<=S} (a/keep)^k where S is the number
    of survived bounces: P(S = k) = keep^k (1 - keep) for k < max_bounce, keep^max_bounce
    for k = max_bounce. Also the standard error factor of the blue sample variance."""
    ks = np.arange(max_bounce + 1)
    pS = keep ** ks * (1 - keep)
    pS[-1] = keep ** max_bounce
    ratio = np.asarray(a, np.float64) / keep
    L = np.array([[np.sum(ratio[c] ** np.arange(k + 1)) for c in range(3)] for k in ks])   # (S, channel)
    mean = (pS[:, None] * L).sum(0)
    var = (pS[:, None] * (L - mean) ** 2).sum(0)
    mu4 = (pS[:, None] * (L - mean) ** 4).sum(0)
    return mean, var, mu4


def roulette_furnace(render, max_bounce=6, W=256, H=256, nsig=5.0):
    """Closed-box furnace with roulette from bounce 0 and one sample per pixel (see module
    doc). Returns (mean result: unbiasedness incl. throughput /= p, blue-variance result:
    the survival probability max(throughput))."""
    v, t = closed_box()
    albedo = np.tile(FURNACE_ALBEDO.astype(np.float32), (len(t), 1))
    emission = np.ones((len(t), 3), np.float32)
    cam = S.camera_lookat((0.1, 0.05, -0.2), (0.3, 0.2, 1.0), 60.0)
    img = render(v, t, albedo, emission, cam, W, H, 1, max_bounce, 0, 17).reshape(-1, 3).astype(np.float64)
    N = W * H
    mean, var, mu4 = furnace_moments(max_bounce)
    exp_mean = np.array([np.sum(FURNACE_ALBEDO[c] ** np.arange(max_bounce + 1)) for c in range(3)])
    assert np.allclose(mean, exp_mean)   # the distribution of S reproduces the unbiased mean
    rm = Result(f"roulette furnace mean (max_bounce {max_bounce})", img.mean(0), exp_mean,
                nsig * np.sqrt(var / N))
    sv = img[:, 2].var(ddof=1)
    rv = Result("roulette furnace blue variance", sv, var[2], nsig * math.sqrt((mu4[2] - var[2] ** 2) / N))
    return rm, rv


NORMALS = [(0.0, 1.0, 0.0), (0.6, 0.8, 0.0), (0.0, 0.8, 0.6), (0.95, 0.31, 0.0), (-0.3, 0.5, -0.81)]


def all_checks(render):
    """Every check above, as a flat list of Results."""
    out = []
    for n in NORMALS:
        out += list(sky_plane(render, n))
    for n in (NORMALS[0], NORMALS[1], NORMALS[3]):
        out.append(ibl_bands(render, n))
    out += list(roulette_furnace(render))
    return out
