"""Seeded synthetic workloads (numpy).

``soup_triangles`` and ``cast_rays`` build the ray-cast workload: one million
triangles with centroids uniform in the unit cube and Gaussian edge vectors,
1920x1080 coherent pinhole rays plus four million incoherent rays.  All
arrays are float32, triangles as (n, 9) rows (v0, v1, v2) and rays as a
(6, n) SoA (ox, oy, oz, dx, dy, dz), the layout the C ABI takes.
"""
from __future__ import annotations

import math

import numpy as np

SOUP_TRIS = 1_000_000
SOUP_SIGMA = 0.003
SOUP_SEED = 2
CAST_WIDTH = 1920
CAST_HEIGHT = 1080
CAST_INCOHERENT = 4_000_000
CAST_SEED = 22


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def soup_triangles(n: int = SOUP_TRIS, seed: int = SOUP_SEED, sigma: float = SOUP_SIGMA) -> np.ndarray:
    """Random triangle soup: centroid c ~ U[0,1)^3, edges e1, e2 ~ N(0, sigma^2)^3,
    v0 = c - (e1 + e2) / 3, v1 = v0 + e1, v2 = v0 + e2.  Returns (n, 9) float32."""
    g = _rng(seed)
    c = g.random((n, 3))
    e1 = g.normal(0.0, sigma, (n, 3))
    e2 = g.normal(0.0, sigma, (n, 3))
    v0 = c - (e1 + e2) / 3.0
    out = np.empty((n, 9), dtype=np.float32)
    out[:, 0:3] = v0
    out[:, 3:6] = v0 + e1
    out[:, 6:9] = v0 + e2
    return out


def pinhole_rays(
    width: int,
    height: int,
    eye=(0.5, 0.5, -1.0),
    fov_deg: float = 40.0,
    jitter: np.ndarray | None = None,
) -> np.ndarray:
    """Pixel-centre pinhole rays looking down +z, vertical field of view
    fov_deg, square pixels, row 0 at the top.  The screen mapping follows
    PinholeCamera::ComputeScaleFromFOV (Cameras/PinholeCamera.cpp:207-212):
    x' = (sx - W/2) * h / H, y' = -(sy - H/2) * h / H, h = 2 tan(fov/2).
    Returns (6, width*height) float32 SoA with unit directions."""
    h = 2.0 * math.tan(math.radians(fov_deg) * 0.5)
    ys, xs = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    sx = xs + 0.5
    sy = ys + 0.5
    if jitter is not None:
        sx = sx + jitter[0].reshape(height, width) - 0.5
        sy = sy + jitter[1].reshape(height, width) - 0.5
    x = (sx - width * 0.5) * h / height
    y = -(sy - height * 0.5) * h / height
    d = np.stack([x.ravel(), y.ravel(), np.ones(width * height)], 0)
    d /= np.linalg.norm(d, axis=0, keepdims=True)
    out = np.empty((6, width * height), dtype=np.float32)
    out[0], out[1], out[2] = eye
    out[3:6] = d
    return out


def incoherent_rays(n: int, seed: int) -> np.ndarray:
    """Origins uniform in the unit cube, directions uniform on the sphere."""
    g = _rng(seed)
    o = g.random((3, n))
    d = g.normal(0.0, 1.0, (3, n))
    d /= np.linalg.norm(d, axis=0, keepdims=True)
    return np.concatenate([o, d], 0).astype(np.float32)


def cast_rays(
    width: int = CAST_WIDTH,
    height: int = CAST_HEIGHT,
    n_incoherent: int = CAST_INCOHERENT,
    seed: int = CAST_SEED,
) -> np.ndarray:
    """The ray-cast ray set: width*height coherent pinhole rays followed by
    n_incoherent incoherent rays, (6, n) float32."""
    a = pinhole_rays(width, height)
    b = incoherent_rays(n_incoherent, seed)
    return np.ascontiguousarray(np.concatenate([a, b], 1))


def rays_from_rows(o: np.ndarray, d: np.ndarray) -> np.ndarray:
    """(n,3) origins and directions to a contiguous (6, n) float32 SoA."""
    return np.ascontiguousarray(np.concatenate([np.asarray(o, np.float32).T, np.asarray(d, np.float32).T], 0))


def near_edge_rays(tris: np.ndarray, n: int, seed: int, band: float = 3e-5) -> np.ndarray:
    """Rays aimed at points within `band` (in barycentrics) of a triangle edge.

    Each ray picks a triangle, an edge of it (u = 0, v = 0 or u + v = 1) and a
    point whose distance from that edge, in barycentric units, is uniform in
    [-band, band]; the origin is uniform in the unit cube.  Near such edges a
    float Moller-Trumbore test of a small triangle seen from afar loses about
    1e-5 in barycentrics, so these rays separate a float pre-filter from the
    double certifier.  Returns (6, n) float32."""
    g = _rng(seed)
    t = np.asarray(tris, np.float64).reshape(-1, 9)
    k = g.integers(0, t.shape[0], n)
    v0, v1, v2 = t[k, 0:3], t[k, 3:6], t[k, 6:9]
    along = g.random(n)
    off = (g.random(n) * 2.0 - 1.0) * band
    edge = g.integers(0, 3, n)
    u = np.where(edge == 0, off, np.where(edge == 1, along, along - off))
    v = np.where(edge == 0, along, np.where(edge == 1, off, 1.0 - along))
    p = v0 + u[:, None] * (v1 - v0) + v[:, None] * (v2 - v0)
    o = g.random((n, 3))
    d = p - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_from_rows(o, d)


def random_object_matrix(g: np.random.Generator, mirror: bool = False, spread: float = 4.0):
    """A random placement: rotation (a reflection too when mirror), non-uniform
    scale in [0.3, 3] and a translation in [-spread, spread]^3, as the
    Matrix4 pair (to_world, to_object) of 16 float64 each, member order
    (row-vector convention)."""
    from .raycast import object_matrices

    q, r = np.linalg.qr(g.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if (np.linalg.det(q) < 0) != mirror:
        q[:, 0] = -q[:, 0]
    lin = np.diag(g.uniform(0.3, 3.0, 3)) @ q
    return object_matrices(lin, g.uniform(-spread, spread, 3))


def instanced_scene(n_meshes: int = 3, n_objects: int = 12, tris_per_mesh: int = 2000, seed: int = 31,
                    n_identity: int = 2, sigma: float = 0.04):
    """A scene of objects: n_meshes seeded soups in their own frame (centroids
    in [-1, 1]^3, edges N(0, (2 sigma)^2)), placed by n_objects objects that
    share them; the first n_identity objects keep identity matrices, one object
    is mirrored.  Returns (meshes, objects) with objects an OBJECT_DTYPE array."""
    from .raycast import make_objects, object_matrices

    g = _rng(seed)
    meshes = []
    for m in range(n_meshes):
        t = soup_triangles(tris_per_mesh, seed=seed * 100 + m, sigma=sigma).astype(np.float64)
        meshes.append(np.ascontiguousarray((t * 2.0 - 1.0).astype(np.float32)))
    entries = []
    for k in range(n_objects):
        mesh = k % n_meshes
        if k < n_identity:
            tw, to = object_matrices()
        else:
            tw, to = random_object_matrix(g, mirror=(k == n_identity))
        entries.append((mesh, tw, to))
    return meshes, make_objects(entries)


def object_rays(meshes, objects, n: int, seed: int, spread: float = 6.0) -> np.ndarray:
    """Rays for a scene of objects: three quarters aimed at world points of
    random triangles of random objects (a quarter of those within 3e-5, in
    barycentrics, of an edge), the rest from random origins in random
    directions; origins uniform in [-spread, spread]^3, directions of random
    length in [0.5, 2] (the reference does not normalize them)."""
    g = _rng(seed)
    o = g.uniform(-spread, spread, (n, 3))
    d = g.normal(size=(n, 3))
    aimed = np.nonzero(g.random(n) < 0.75)[0]
    k = aimed.size
    which = g.integers(0, len(objects), k)
    pick = g.random(k)
    u, v = g.random(k), g.random(k)
    flip = u + v > 1.0
    u, v = np.where(flip, 1.0 - u, u), np.where(flip, 1.0 - v, v)
    along, off, edge = g.random(k), (g.random(k) * 2.0 - 1.0) * 3e-5, g.integers(0, 3, k)
    eu = np.where(edge == 0, off, np.where(edge == 1, along, along - off))
    ev = np.where(edge == 0, along, np.where(edge == 1, off, 1.0 - along))
    near = pick < 0.25
    u, v = np.where(near, eu, u), np.where(near, ev, v)
    tri_pick = g.random(k)
    for ob_i in range(len(objects)):
        sel = which == ob_i
        if not sel.any():
            continue
        ob = objects[ob_i]
        t = np.asarray(meshes[int(ob["mesh"])], np.float64).reshape(-1, 9)
        tri = t[np.minimum((tri_pick[sel] * t.shape[0]).astype(np.int64), t.shape[0] - 1)]
        p = tri[:, 0:3] + u[sel, None] * (tri[:, 3:6] - tri[:, 0:3]) + v[sel, None] * (tri[:, 6:9] - tri[:, 0:3])
        m = np.asarray(ob["to_world"], np.float64).reshape(4, 4)
        d[aimed[sel]] = p @ m[:3, :3] + m[3, :3] - o[aimed[sel]]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= g.uniform(0.5, 2.0, (n, 1))
    return rays_from_rows(o, d)
