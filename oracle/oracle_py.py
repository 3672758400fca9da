"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product never imports it.

Scenes are (pos (nv,3) float32, idx (nt,3) uint32); trees are the reference's
bbvh_node records (numpy structured array, 56 B each) with a root and the
leaf-slot index; rays are (8, n) float32 planes ox, oy, oz, tmin, dx, dy, dz,
tmax.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
FLT_MAX = np.float32(3.4028234663852886e38)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run make in oracle/")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_max_threads.restype = c_int
        L.orc_time_closest.restype = c_double
        _lib = L
    return _lib


def _p(a):
    return c_void_p(a.ctypes.data) if a is not None else None


def _planes(rays):
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    arr = (c_void_p * 8)(*[rays[k].ctypes.data for k in range(8)])
    return rays, arr


def _scene(pos, idx):
    return (np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3),
            np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3))


def trace_closest(pos, idx, nodes, root, index, rays, row=0, threads=0, counts=False):
    """binary_bvh_tracer::closest_hit over every ray. Returns t, u, v, id[, counts (n,2)]."""
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    t = np.empty(n, np.float32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    i = np.empty(n, np.uint32)
    c = np.zeros((n, 2), np.uint32) if counts else None
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    lib().orc_trace_closest(_p(pos), _p(idx), c_uint32(len(idx)), _p(nodes), c_uint32(len(nodes)), c_uint32(root),
                            _p(index), planes, c_uint32(n), c_uint32(row), _p(t), _p(u), _p(v), _p(i), _p(c),
                            c_int(threads))
    return (t, u, v, i, c) if counts else (t, u, v, i)


def trace_any(pos, idx, nodes, root, index, rays, threads=0):
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    occ = np.empty(n, np.uint8)
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    lib().orc_trace_any(_p(pos), _p(idx), c_uint32(len(idx)), _p(nodes), c_uint32(len(nodes)), c_uint32(root),
                        _p(index), planes, c_uint32(n), _p(occ), c_int(threads))
    return occ


class Alpha:
    """Opacity inputs of the alpha-aware tracers: per-vertex texture
    coordinates (n_vert, 2), per-triangle material ids (or None), and a dict
    material -> RGBA texture (h, w, 4)."""

    def __init__(self, tc, mat, textures, n_mats):
        self.tc = np.ascontiguousarray(tc, dtype=np.float32).reshape(-1, 2)
        self.mat = None if mat is None else np.ascontiguousarray(mat, dtype=np.uint32)
        self.mat_tex = np.full(max(n_mats, 1), -1, np.int32)
        tex, texels, first = [], [], 0
        for k, (m, img) in enumerate(sorted(textures.items())):
            img = np.ascontiguousarray(img, dtype=np.float32)
            h, w = img.shape[:2]
            self.mat_tex[m] = k
            tex.append((first, w, h))
            texels.append(img.reshape(-1, 4))
            first += w * h
        self.tex = np.array(tex if tex else [(0, 1, 1)], np.uint32).reshape(-1, 3)
        self.texels = np.concatenate(texels) if texels else np.zeros((1, 4), np.float32)
        self.n_mats = n_mats

    def args(self):
        return (_p(self.tc), _p(self.mat), _p(self.mat_tex), c_uint32(self.n_mats), _p(self.tex), _p(self.texels))


def trace_closest_alpha(pos, idx, nodes, root, index, rays, alpha, threads=0):
    """binary_bvh_tracer<..., alpha_aware>::closest_hit. Returns t, u, v, id."""
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    t = np.empty(n, np.float32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    i = np.empty(n, np.uint32)
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    lib().orc_trace_closest_alpha(_p(pos), _p(idx), c_uint32(len(idx)), _p(nodes), c_uint32(len(nodes)),
                                  c_uint32(root), _p(index), *alpha.args(), planes, c_uint32(n), _p(t), _p(u), _p(v),
                                  _p(i), c_int(threads))
    return t, u, v, i


def trace_any_alpha(pos, idx, nodes, root, index, rays, alpha, threads=0):
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    occ = np.empty(n, np.uint8)
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    lib().orc_trace_any_alpha(_p(pos), _p(idx), c_uint32(len(idx)), _p(nodes), c_uint32(len(nodes)), c_uint32(root),
                              _p(index), *alpha.args(), planes, c_uint32(n), _p(occ), c_int(threads))
    return occ


def opacity(pos, idx, alpha, tri, beta, gamma):
    """diff_geom::opacity of triangle tri[i] at (beta[i], gamma[i])."""
    pos, idx = _scene(pos, idx)
    tri = np.ascontiguousarray(tri, dtype=np.uint32)
    beta = np.ascontiguousarray(beta, dtype=np.float32)
    gamma = np.ascontiguousarray(gamma, dtype=np.float32)
    out = np.empty(len(tri), np.float32)
    lib().orc_opacity(_p(pos), _p(idx), c_uint32(len(idx)), *alpha.args(), _p(tri), _p(beta), _p(gamma),
                      c_uint32(len(tri)), _p(out))
    return out


def seq_closest(pos, idx, rays, threads=0):
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    t = np.empty(n, np.float32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    i = np.empty(n, np.uint32)
    lib().orc_seq_closest(_p(pos), _p(idx), c_uint32(len(idx)), planes, c_uint32(n), _p(t), _p(u), _p(v), _p(i),
                          c_int(threads))
    return t, u, v, i


def seq_any(pos, idx, rays, threads=0):
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    occ = np.empty(n, np.uint8)
    lib().orc_seq_any(_p(pos), _p(idx), c_uint32(len(idx)), planes, c_uint32(n), _p(occ), c_int(threads))
    return occ


def fp64_intersect(pos, idx, tri, rays):
    """Double-precision intersection of ray i with triangle tri[i]: t, beta, gamma, inside."""
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    tri = np.ascontiguousarray(tri, dtype=np.uint32)
    t = np.empty(n, np.float64)
    b = np.empty(n, np.float64)
    g = np.empty(n, np.float64)
    ins = np.empty(n, np.uint8)
    lib().orc_fp64_intersect(_p(pos), _p(idx), _p(tri), planes, c_uint32(n), _p(t), _p(b), _p(g), _p(ins))
    return t, b, g, ins


def camera_rays(pos, dir, up, fovy, w, h, ox=None, oy=None):
    """Camera rays (8, w*h); ox, oy: optional sub-pixel offsets per pixel."""
    out = np.empty((8, w * h), np.float32)
    arr = (c_void_p * 8)(*[out[k].ctypes.data for k in range(8)])
    P = np.asarray(pos, np.float32)
    D = np.asarray(dir, np.float32)
    U = np.asarray(up, np.float32)
    ox = None if ox is None else np.ascontiguousarray(ox, np.float32)
    oy = None if oy is None else np.ascontiguousarray(oy, np.float32)
    lib().orc_camera_rays(_p(P), _p(D), _p(U), c_float(fovy), c_int(w), c_int(h), _p(ox), _p(oy), arr)
    return out


def incoherent_rays(sphere1, sphere2, r1, r2, r_max, seed, n):
    """orc_incoherent_rays: (8, n) rays between two balls."""
    out = np.empty((8, n), np.float32)
    arr = (c_void_p * 8)(*[out[k].ctypes.data for k in range(8)])
    s1 = np.asarray(sphere1, np.float32)
    s2 = np.asarray(sphere2, np.float32)
    lib().orc_incoherent_rays(_p(s1), _p(s2), c_float(r1), c_float(r2), c_float(r_max), c_uint64(seed), c_uint32(n),
                              arr)
    return out


def cosine_rays(pos, idx, in_rays, t, u, v, ids, xi0, xi1, normals=None):
    """orc_cosine_rays; normals: optional (n_vert, 3) shading normals."""
    pos, idx = _scene(pos, idx)
    vn = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    in_rays, inp = _planes(in_rays)
    n = in_rays.shape[1]
    out = np.empty((8, n), np.float32)
    arr = (c_void_p * 8)(*[out[k].ctypes.data for k in range(8)])
    t = np.ascontiguousarray(t, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    ids = np.ascontiguousarray(ids, np.uint32)
    xi0 = np.ascontiguousarray(xi0, np.float32)
    xi1 = np.ascontiguousarray(xi1, np.float32)
    lib().orc_cosine_rays(_p(pos), _p(idx), _p(vn), inp, _p(t), _p(u), _p(v), _p(ids), _p(xi0), _p(xi1), c_uint32(n),
                          arr)
    return out


def time_closest(pos, idx, nodes, root, index, rays, row=0, threads=0):
    """Seconds of one batch closest-hit call (OpenMP over rows)."""
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    t = np.empty(n, np.float32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    i = np.empty(n, np.uint32)
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    return lib().orc_time_closest(_p(pos), _p(idx), c_uint32(len(idx)), _p(nodes), c_uint32(len(nodes)),
                                  c_uint32(root), _p(index), planes, c_uint32(n), c_uint32(row), _p(t), _p(u), _p(v),
                                  _p(i), c_int(threads))


def max_threads():
    return lib().orc_max_threads()


def adjoint_fp64(pos, idx, tri, rays, gt, gb, gg, fd_step=None):
    """FP64 gradient of sum(gt*t + gb*beta + gg*gamma) per ray w.r.t. (A, B, C, o, d):
    analytic (fd_step None) or by central differences with step fd_step.  (n, 15)."""
    pos, idx = _scene(pos, idx)
    rays, planes = _planes(rays)
    n = rays.shape[1]
    tri = np.ascontiguousarray(tri, dtype=np.uint32)
    g = [np.ascontiguousarray(x, dtype=np.float64) for x in (gt, gb, gg)]
    out = np.empty((n, 15), np.float64)
    if fd_step is None:
        lib().orc_adjoint_fp64(_p(pos), _p(idx), _p(tri), planes, c_uint32(n), _p(g[0]), _p(g[1]), _p(g[2]),
                               _p(out))
    else:
        lib().orc_adjoint_fd(_p(pos), _p(idx), _p(tri), planes, c_uint32(n), _p(g[0]), _p(g[1]), _p(g[2]),
                             c_double(fd_step), _p(out))
    return out


def pcg_sequence(initstate, initseq, n, raw=True):
    """First n outputs of PCG32 seeded with (initstate, initseq): uint32 values
    (raw) or next_float values."""
    out = np.empty(n, np.uint32)
    lib().orc_pcg_sequence(ctypes.c_uint64(initstate), ctypes.c_uint64(initseq), c_uint32(n), c_int(1 if raw else 0),
                           _p(out))
    return out if raw else out.view(np.float32)


def frame_seeds(seed, first_frame, n_frames):
    """Frame seeds of render frames first_frame .. first_frame + n_frames - 1:
    seed * 0x100000001B3 + frame (mod 2^64)."""
    base = (int(seed) * 0x100000001B3) % (1 << 64)
    return np.array([(base + first_frame + k) % (1 << 64) for k in range(n_frames)], np.uint64)


def path_trace(pos, idx, mat, materials, nodes, root, index, cam, seeds, max_depth=4, rr_start=3, rr_survive=0.8,
               emitters_reflect=False, normals=None, jitter=True, threads=0):
    """The workload's diffuse path tracer over the frame seeds `seeds`, in
    order.  materials: (n_mat, 8) albedo rgb, 0, emission rgb, 0; mat: material
    per triangle (None: material 0).  Returns image (h, w, 3) accumulated over
    the frames, rays traced per level, and a per-pixel signature of the last
    frame's hit sequence."""
    pos, idx = _scene(pos, idx)
    cam = np.ascontiguousarray(cam[:12], dtype=np.float32)
    w, h = int(cam[10]), int(cam[11])
    mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, 8)
    mat = None if mat is None else np.ascontiguousarray(mat, dtype=np.uint32)
    vn = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    img = np.zeros((h, w, 3), np.float32)
    levels = np.zeros(max_depth + 1, np.uint64)
    sig = np.zeros(w * h, np.uint64)
    lib().orc_path_trace(_p(pos), _p(idx), c_uint32(len(idx)), _p(mat), _p(mats), _p(vn), _p(nodes),
                         c_uint32(len(nodes)), c_uint32(root), _p(index), _p(cam), c_int(max_depth), c_int(rr_start),
                         c_float(rr_survive), c_int(1 if emitters_reflect else 0), c_int(1 if jitter else 0),
                         _p(seeds), c_uint32(len(seeds)),
                         _p(img), _p(levels), _p(sig), c_int(threads))
    return img, levels, sig.reshape(h, w)


DIRECT_MODES = {"uniform": 0, "cosine": 1, "light": 2}


def direct(pos, idx, mat, materials, nodes, root, index, cam, mode, seeds, points=None, threads=0, normals=None):
    """The wavefront direct-lighting algorithm (orc_direct) in mode "uniform",
    "cosine" or "light", one pass per seed.  points: (k, 6) pos + col of point
    lights; normals: optional (n_vert, 3) shading normals.  Returns the
    framebuffer (h, w, 4): rgb sums and weights."""
    pos, idx = _scene(pos, idx)
    vn = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    cam = np.ascontiguousarray(cam[:12], dtype=np.float32)
    w, h = int(cam[10]), int(cam[11])
    mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, 8)
    mat = None if mat is None else np.ascontiguousarray(mat, dtype=np.uint32)
    pts = np.zeros((0, 6), np.float32) if points is None else np.ascontiguousarray(points, np.float32).reshape(-1, 6)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    nodes = np.ascontiguousarray(nodes)
    index = np.ascontiguousarray(index, dtype=np.uint32)
    fb = np.zeros((h, w, 4), np.float32)
    lib().orc_direct(_p(pos), _p(idx), c_uint32(len(idx)), _p(mat), _p(mats), _p(vn), _p(nodes), c_uint32(len(nodes)),
                     c_uint32(root), _p(index), _p(cam), c_int(DIRECT_MODES[mode]), _p(pts), c_uint32(len(pts)),
                     _p(seeds), c_uint32(len(seeds)), _p(fb), c_int(threads))
    return fb
