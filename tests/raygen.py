"""Seeded ray sets for parity tests: random, camera, and adversarial rays aimed
exactly at triangle vertices and edge midpoints (shared edges and vertices
are where watertightness and the tie rule matter)."""
import numpy as np

from solhip.records import RAY


def world_tris(sd, max_tris=None, rng=None):
    """world-space triangles (n, 3, 3) of the scene (float64 transform)"""
    out = []
    for i, d in enumerate(sd.instance_descs):
        b = sd.blas[int(d["blas"])]
        P = b.vertices["pos"][:, :3].astype(np.float64)[b.indices.reshape(-1, 3)]
        M = d["transform"].reshape(3, 4).astype(np.float64)
        W = P @ M[:, :3].T + M[:, 3]
        out.append(W)
    T = np.concatenate(out) if out else np.zeros((0, 3, 3))
    if max_tris is not None and len(T) > max_tris:
        rng = rng or np.random.default_rng(0)
        T = T[rng.choice(len(T), max_tris, replace=False)]
    return T


def random_rays(n, center, radius, seed=0, tmin=0.0, tmax=1e4):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY)
    r["o"] = np.asarray(center) + rng.uniform(-radius, radius, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def aimed_rays(sd, n, seed=1, dist=3.0):
    """rays from random points toward exact vertices and edge midpoints"""
    rng = np.random.default_rng(seed)
    T = world_tris(sd, max_tris=max(1, n // 4), rng=rng)
    if len(T) == 0:
        return np.zeros(0, RAY)
    targets = []
    for t in T:
        targets += [t[0], t[1], t[2], 0.5 * (t[0] + t[1]), 0.5 * (t[1] + t[2]), (t[0] + t[1] + t[2]) / 3.0]
    targets = np.asarray(targets, np.float32)[:n]
    m = len(targets)
    dirs = rng.normal(size=(m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    o = (targets.astype(np.float64) - dist * dirs).astype(np.float32)
    r = np.zeros(m, RAY)
    r["o"] = o
    r["d"] = (targets.astype(np.float64) - o.astype(np.float64)).astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = 1e4
    return r


def camera_rays(sd, W, H, oracle):
    u = sd.uniforms(W, H, 0)
    r = np.zeros(W * H, RAY)
    k = 0
    for y in range(H):
        for x in range(W):
            o, d = oracle.camera_ray(u, x, y, W, H, 0.5, 0.5)
            r[k]["o"] = o
            r[k]["d"] = d
            k += 1
    r["tmin"] = 1e-3
    r["tmax"] = 1e4
    return r


def axis_rays(n, seed=3, span=1.5):
    """rays parallel to the coordinate axes (zero direction components)"""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY)
    r["o"] = rng.uniform(-span, span, size=(n, 3))
    ax = rng.integers(0, 3, size=n)
    sg = rng.choice([-1.0, 1.0], size=n)
    d = np.zeros((n, 3))
    d[np.arange(n), ax] = sg
    r["o"][np.arange(n), ax] = -3.0 * sg
    r["d"] = d
    r["tmin"] = 0.0
    r["tmax"] = 1e4
    return r
