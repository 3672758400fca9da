"""Synthetic multi-scene workload: many random-triangle scenes of uneven size,
each with its own rays, traced in one descriptor-table launch.

Scene k has n_tris ~ logU[tri_lo, tri_hi] triangles from the random-triangle
generator (centroids U[0,1]^3, vertices N(c, sigma^2) with sigma = 0.02 *
(8192 / n_tris)^(1/3), so that the expected triangle area per unit volume
stays that of the 8,192-triangle scene) and n_rays ~ logU[ray_lo, ray_hi]
rays: half pinhole camera rays (pos (0.5,0.5,-1), looking down +z, vertical
fovy 60 degrees, 4:3), half rays from origins U[0,1]^3 in uniform directions.
"""
import concurrent.futures as cf
import os

import numpy as np

import hiprt

FLT_MAX = np.float32(3.4028234663852886e38)


def sphere_rays(n, rng, tmin=1e-4):
    o = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    z = 1.0 - 2.0 * rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    rays = np.empty((8, n), np.float32)
    rays[0:3] = o.T
    rays[3] = tmin
    rays[4:7] = d.T
    rays[7] = FLT_MAX
    return rays


def scene_sizes(n_scenes, seed, tri_lo=1000, tri_hi=500_000, ray_lo=4000, ray_hi=2_000_000):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_scenes):
        nt = int(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        nr = int(np.exp(rng.uniform(np.log(ray_lo), np.log(ray_hi))))
        out.append((nt, nr))
    return out


def make_scene(k, n_tris, n_rays, seed):
    """Host scene, BVH (SAH 16 planes, 4 tris/leaf, ESC) and rays of scene k."""
    sigma = 0.02 * (8192.0 / n_tris) ** (1.0 / 3.0)
    sc = hiprt.HostScene.random(n_tris, sigma, seed * 1000 + k)
    bvh = hiprt.Bvh(sc)
    w = max(1, int(np.sqrt(n_rays / 2 * 4 / 3)))
    h = max(1, (n_rays // 2) // w)
    cam = np.array([0.5, 0.5, -1, 0, 0, 1, 0, 1, 0, 60, w, h], np.float32)
    rng = np.random.default_rng(seed * 1000 + 500 + k)
    rays = np.concatenate([hiprt.camera_rays(cam), sphere_rays(n_rays - w * h, rng)], 1)
    return sc, bvh, rays


def multi_scene(n_scenes=64, seed=4, workers=None, **sizes):
    """[(HostScene, Bvh, rays (8, n) float32)] for every scene.  The builds
    run in threads: the library's calls release the interpreter lock."""
    dims = scene_sizes(n_scenes, seed, **sizes)
    workers = workers or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = [ex.submit(make_scene, k, nt, nr, seed) for k, (nt, nr) in enumerate(dims)]
        return [f.result() for f in futs]
