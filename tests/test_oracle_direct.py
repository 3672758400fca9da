"""Pins of the oracle's wavefront direct lighting (oracle.cpp: orc_direct).

A floor (y = 0, albedo a) seen from above, under emitters:
* cosine mode, five emitting faces around it: every sample is e * a exactly
  up to rounding (cos / pdf cancels: radiance = e a/pi cos / (cos/pi));
* uniform mode: 2 e a cos, with cos = the first uniform of the pixel's
  PCG32 stream (z of the uniform hemisphere sample);
* light mode over one emitting triangle, two of different power, and a point
  light: the per-sample estimate e a/pi cos_x cos_l area / dist^2 / p_light
  recomputed in float64 from the same uniforms (light index by the power
  distribution's running sums, point by uniform_sample_triangle);
* the reference's quirks: the integration uses the floor's normal unflipped,
  so a floor whose winding faces down gets no light; a light facing away
  gives none; a camera hit on an emitter adds (e, 1) in the uniform and
  cosine modes (then (0, 1) from the integration) but (e, 0) in light mode;
  a miss adds (0, 1).
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from pcg import Pcg32, stream_id

CAM = np.array([0.41, 0.8, 0.52, 0, -1, 0, 0, 0, 1, 50, 24, 24], np.float32)
SEEDS = orc.frame_seeds(11, 0, 2)


def _floor(up=True):
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    idx = [(0, 2, 1), (0, 3, 2)] if up else [(0, 1, 2), (0, 2, 3)]
    return P, idx


def _scene(parts):
    """parts: [(positions (k,3), triangles [(i,j,l)], material)] -> pos, idx, mat."""
    pos, idx, mat = [], [], []
    base = 0
    for P, tris, m in parts:
        pos.append(np.asarray(P, np.float32))
        idx += [(a + base, b + base, c + base) for a, b, c in tris]
        mat += [m] * len(tris)
        base += len(P)
    return np.concatenate(pos), np.array(idx, np.uint32), np.array(mat, np.uint32)


def _walls():
    """The four walls and the ceiling of the unit box (any winding)."""
    P = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)
    v = lambda x, y, z: 4 * x + 2 * y + z
    quads = [(v(0, 1, 0), v(0, 1, 1), v(1, 1, 1), v(1, 1, 0)), (v(0, 0, 0), v(0, 0, 1), v(0, 1, 1), v(0, 1, 0)),
             (v(1, 0, 0), v(1, 1, 0), v(1, 1, 1), v(1, 0, 1)), (v(0, 0, 0), v(0, 1, 0), v(1, 1, 0), v(1, 0, 0)),
             (v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1))]
    tris = []
    for a, b, c, d in quads:
        tris += [(a, b, c), (a, c, d)]
    return P, tris


def _run(pos, idx, mat, mats, mode, points=None, cam=CAM, seeds=SEEDS, normals=None):
    sc = hiprt.HostScene.from_arrays(pos, idx, mat)
    bvh = hiprt.Bvh(sc)
    return orc.direct(pos, idx, mat, mats, bvh.nodes, bvh.root, bvh.index, cam, mode, seeds, points, normals=normals)


A = 0.5
E = np.array([2.0, 3.0, 4.0])
MATS = np.array([[A, A, A, 0, 0, 0, 0, 0], [0.3, 0.3, 0.3, 0, *E, 0]], np.float32)


def _floor_points(cam=CAM):
    """Where each pixel's camera ray meets y = 0 (float64)."""
    rays = orc.camera_rays(cam[0:3], cam[3:6], cam[6:9], float(cam[9]), int(cam[10]), int(cam[11]))
    o, d = rays[0:3].astype(np.float64), rays[4:7].astype(np.float64)
    return (o - d * (o[1] / d[1])).T


def _uniforms(seed, n, k):
    g = Pcg32(stream_id(int(seed), np.arange(n, dtype=np.uint64), 0))
    return [g.next_float().astype(np.float64) for _ in range(k)]


def test_cosine_mode_cancels_to_albedo_times_emission():
    F, ft = _floor(up=True)
    W, wt = _walls()
    pos, idx, mat = _scene([(F, ft, 0), (W, wt, 1)])
    fb = _run(pos, idx, mat, MATS, "cosine").reshape(-1, 4)
    assert np.all(fb[:, 3] == len(SEEDS))
    np.testing.assert_allclose(fb[:, :3] / len(SEEDS), np.tile(A * E, (len(fb), 1)), rtol=2e-6)


def test_uniform_mode_is_twice_the_cosine_of_the_sample():
    F, ft = _floor(up=True)
    W, wt = _walls()
    pos, idx, mat = _scene([(F, ft, 0), (W, wt, 1)])
    fb = _run(pos, idx, mat, MATS, "uniform").reshape(-1, 4)
    n = len(fb)
    want = sum(2 * A * E[None, :] * _uniforms(s, n, 1)[0][:, None] for s in SEEDS)
    np.testing.assert_allclose(fb[:, :3], want, rtol=1e-5, atol=1e-6)


def test_floor_facing_down_gets_no_light():
    # the integration steps test the direction against the hit's own normal,
    # unflipped (rt/cpu/bounce.cpp:108-151)
    F, ft = _floor(up=False)
    W, wt = _walls()
    pos, idx, mat = _scene([(F, ft, 0), (W, wt, 1)])
    for mode in ("uniform", "cosine"):
        fb = _run(pos, idx, mat, MATS, mode).reshape(-1, 4)
        assert not fb[:, :3].any() and np.all(fb[:, 3] == len(SEEDS))


def _light_tri(y, big, down=True):
    # both at the same height never occlude each other
    P = np.array([[-0.5, y, -0.5], [1.5, y, -0.5], [-0.5, y, 1.5]], np.float32) if big else \
        np.array([[1.1, y, 0.2], [1.8, y, 0.2], [1.1, y, 0.9]], np.float32)
    return P, [(0, 1, 2)] if down else [(0, 2, 1)]


def _tri_estimate(x, L, e, xi0, xi1, p_light, light_normal=None):
    """e a/pi cos_x cos_l area / dist^2 / p_light for a point of triangle L
    (cos_l against light_normal when given, else the geometric normal)."""
    a, b, c = (np.asarray(p, np.float64) for p in L)
    su0 = np.sqrt(xi0)
    bx, by = 1 - su0, xi1 * su0
    target = (1 - bx - by)[:, None] * a + bx[:, None] * b + by[:, None] * c
    nl = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(nl)
    nl /= np.linalg.norm(nl)
    if light_normal is not None:
        nl = np.asarray(light_normal, np.float64)
    w = target - x
    dist = np.linalg.norm(w, axis=1)
    w /= dist[:, None]
    cos_x, cos_l = w[:, 1], -(w @ nl)
    est = e[None, :] * (A / np.pi) * (cos_x * cos_l * area / dist ** 2 / p_light)[:, None]
    return np.where((cos_l > 0)[:, None], est, 0.0)


def test_light_mode_one_triangle():
    F, ft = _floor(up=True)
    L, lt = _light_tri(1.0, big=True)
    pos, idx, mat = _scene([(F, ft, 0), (L, lt, 1)])
    fb = _run(pos, idx, mat, MATS, "light").reshape(-1, 4)
    x = _floor_points()
    want = 0.0
    for s in SEEDS:
        _, xi0, xi1 = _uniforms(s, len(x), 3)
        want = want + _tri_estimate(x, L, E, xi0, xi1, 1.0)
    assert np.all(fb[:, 3] == len(SEEDS))
    np.testing.assert_allclose(fb[:, :3], want, rtol=2e-5, atol=1e-7)


def test_cosine_mode_weighs_by_the_shading_normal():
    """With the floor's normals tilted by 0.5 rad from its geometric normal,
    each cosine sample w (drawn about the geometric normal, pdf cos_g / pi)
    adds e * a * max(0, w . ns) / cos_g: integrate_dir_sample weighs by the
    shading normal (rt/cpu/bounce.cpp:108-130)."""
    F, ft = _floor(up=True)
    W, wt = _walls()
    pos, idx, mat = _scene([(F, ft, 0), (W, wt, 1)])
    tilt = np.array([np.sin(0.5), np.cos(0.5), 0.0])
    normals = np.zeros((len(pos), 3), np.float32)
    normals[:4] = tilt
    normals[4:] = [0, 1, 0]   # the emitting walls: only their emission counts
    fb = _run(pos, idx, mat, MATS, "cosine", normals=normals).reshape(-1, 4)
    n = len(fb)
    want = 0.0
    for s in SEEDS:
        xi0, xi1 = _uniforms(s, n, 2)
        r, th = np.sqrt(xi0), 2 * np.pi * xi1
        dx, dy = r * np.cos(th), r * np.sin(th)
        z = np.sqrt(np.maximum(0.0, 1 - dx * dx - dy * dy))
        w = np.stack([-dx, z, dy], 1)   # align(sample, +y)
        c = w @ tilt
        want = want + A * E[None, :] * (np.maximum(c, 0) / z)[:, None]
    assert np.all(fb[:, 3] == len(SEEDS))
    np.testing.assert_allclose(fb[:, :3], want, rtol=1e-4, atol=1e-6)


def test_light_mode_uses_the_lights_interpolated_normal():
    """trianglelight::sample_Li takes the light's vertex normals interpolated
    at the sampled point for cos_light (libgi/scene.cpp:534-559): tilting
    them changes the estimate accordingly."""
    F, ft = _floor(up=True)
    L, lt = _light_tri(1.0, big=True)
    pos, idx, mat = _scene([(F, ft, 0), (L, lt, 1)])
    nl = np.array([np.sin(0.4), -np.cos(0.4), 0.0])
    normals = np.zeros((len(pos), 3), np.float32)
    normals[:4] = [0, 1, 0]   # the floor's shading normal = its geometric normal
    normals[4:] = nl
    fb = _run(pos, idx, mat, MATS, "light", normals=normals).reshape(-1, 4)
    x = _floor_points()
    want = 0.0
    for s in SEEDS:
        _, xi0, xi1 = _uniforms(s, len(x), 3)
        want = want + _tri_estimate(x, L, E, xi0, xi1, 1.0, light_normal=nl)
    assert np.all(fb[:, 3] == len(SEEDS))
    np.testing.assert_allclose(fb[:, :3], want, rtol=5e-5, atol=1e-7)
    plain = _run(pos, idx, mat, MATS, "light").reshape(-1, 4)
    assert not np.allclose(plain[:, :3], fb[:, :3], rtol=1e-3)


def test_light_mode_light_facing_away_and_floor_facing_down():
    for floor_up, light_down in ((True, False), (False, True)):
        F, ft = _floor(up=floor_up)
        L, lt = _light_tri(1.0, big=True, down=light_down)
        pos, idx, mat = _scene([(F, ft, 0), (L, lt, 1)])
        fb = _run(pos, idx, mat, MATS, "light").reshape(-1, 4)
        assert not fb[:, :3].any() and np.all(fb[:, 3] == len(SEEDS))


def test_light_mode_picks_lights_by_power():
    F, ft = _floor(up=True)
    L1, t1 = _light_tri(1.0, big=True)
    L2, t2 = _light_tri(1.0, big=False)
    e2 = np.array([9.0, 1.0, 0.5])
    mats = np.concatenate([MATS, np.array([[0.3, 0.3, 0.3, 0, *e2, 0]], np.float32)])
    pos, idx, mat = _scene([(F, ft, 0), (L1, t1, 1), (L2, t2, 2)])
    fb = _run(pos, idx, mat, mats, "light").reshape(-1, 4)

    def power(L, e):
        a, b, c = (np.asarray(p, np.float32) for p in L)
        area2 = np.float32(np.linalg.norm(np.cross(b - a, c - a)))
        rgb = e.astype(np.float32) * np.float32(0.5) * area2 * np.float32(np.pi)
        return float(np.float32(0.212671) * rgb[0] + np.float32(0.715160) * rgb[1] + np.float32(0.072169) * rgb[2])

    p1, p2 = power(L1, E), power(L2, e2)
    c1 = p1 / (p1 + p2)
    x = _floor_points()
    want = 0.0
    picks = []
    for s in SEEDS:
        xl, xi0, xi1 = _uniforms(s, len(x), 3)
        second = xl > c1   # lower_bound over (0, c1, 1), minus one
        picks.append(second)
        est1 = _tri_estimate(x, L1, E, xi0, xi1, p1 / (p1 + p2))
        est2 = _tri_estimate(x, L2, e2, xi0, xi1, p2 / (p1 + p2))
        want = want + np.where(second[:, None], est2, est1)
    assert 0.05 < np.mean(picks) < 0.95   # both lights are sampled
    np.testing.assert_allclose(fb[:, :3], want, rtol=5e-5, atol=1e-6)


def test_light_mode_point_light():
    F, ft = _floor(up=True)
    pos, idx, mat = _scene([(F, ft, 0)])
    col = np.array([3.0, 2.0, 1.0])
    lp = np.array([0.5, 0.9, 0.45])
    fb = _run(pos, idx, mat, MATS[:1], "light", points=[[*lp, *col]]).reshape(-1, 4)
    x = _floor_points()
    w = lp[None, :] - x
    dist = np.linalg.norm(w, axis=1)
    cos_x = w[:, 1] / dist
    want = len(SEEDS) * col[None, :] / dist[:, None] ** 2 * (A / np.pi) * cos_x[:, None]
    np.testing.assert_allclose(fb[:, :3], want, rtol=2e-5)


def test_emitter_and_miss_weights():
    F, ft = _floor(up=True)
    pos, idx, mat = _scene([(F, ft, 1)])   # the floor itself emits
    L, lt = _light_tri(5.0, big=True)
    pos2, idx2, mat2 = _scene([(F, ft, 1), (L, lt, 1)])
    S = len(SEEDS)
    for mode, w in (("uniform", 2), ("cosine", 2)):
        fb = _run(pos, idx, mat, MATS, mode).reshape(-1, 4)
        assert np.all(fb[:, 3] == w * S)
        np.testing.assert_allclose(fb[:, :3], np.tile(S * E, (len(fb), 1)), rtol=1e-6)
    fb = _run(pos2, idx2, mat2, MATS, "light").reshape(-1, 4)
    assert np.all(fb[:, 3] == S)
    np.testing.assert_allclose(fb[:, :3], np.tile(S * E, (len(fb), 1)), rtol=1e-6)
    # a camera looking away from everything: (0, 1) per sample
    away = CAM.copy()
    away[1], away[4] = 6.0, 1.0
    for mode in ("uniform", "cosine", "light"):
        fb = _run(pos2, idx2, mat2, MATS, mode, cam=away).reshape(-1, 4)
        assert not fb[:, :3].any() and np.all(fb[:, 3] == S)
