"""Pins for the oracle's alpha-aware traversal (binary_bvh_tracer with
alpha_aware = true, rt/cpu/bvh2.cpp:109-119, 171-176) and its texture lookup
(texture2d::sample, libgi/scene.h:28-36; diff_geom's texture coordinate,
libgi/rt.cpp:11).

* lookup: with vertex coordinates (0,0), (1,0), (0,1) the coordinate is
  exactly (beta, gamma), so dyadic barycentrics name texels in closed form:
  texel x = floor(u w + 1/2) with u = w wraps to 0, integer offsets of the
  coordinates wrap away, half-texel points round up;
* threshold: alpha exactly 0.5 is transparent, anything above opaque;
* traversal: an all-opaque texture changes nothing (bit-identical to the
  plain traversal); a fully transparent texture on one material equals brute
  force over the other material's triangles alone, closest and any hit.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, uniform_sphere_rays

W, H = 8, 4


def _lookup_scene(offset=(0.0, 0.0)):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    tc = np.array([[0, 0], [1, 0], [0, 1]], np.float32) + np.float32(offset)
    img = np.zeros((H, W, 4), np.float32)
    img[..., 3] = (np.arange(H * W, dtype=np.float32).reshape(H, W) + 1) / 64.0   # texel (x, y) -> (y W + x + 1) / 64
    return pos, idx, orc.Alpha(tc, None, {0: img}, 1)


def _texel_alpha(x, y):
    return np.float32((y * W + x + 1) / 64.0)


@pytest.mark.parametrize("offset", [(0.0, 0.0), (3.0, -2.0), (-5.0, 7.0)])
def test_lookup_names_texels_in_closed_form(offset):
    pos, idx, a = _lookup_scene(offset)
    bs, gs, want = [], [], []
    for x in range(W + 1):
        for y in range(H + 1):
            b, g = x / W, y / H
            if b + g > 1.0:
                continue
            bs.append(b)
            gs.append(g)
            want.append(_texel_alpha(x % W, y % H))   # u = 1 wraps to texel 0
    got = orc.opacity(pos, idx, a, np.zeros(len(bs), np.uint32), bs, gs)
    assert np.array_equal(got, np.array(want, np.float32))


def test_lookup_rounds_half_texels_up():
    pos, idx, a = _lookup_scene()
    # u w + 1/2 truncated: just below a half-texel stays, the half-texel point moves up
    b = np.array([(2 + 0.5) / W - 2.0 ** -20, (2 + 0.5) / W, 0.5 / W - 2.0 ** -20], np.float32)
    g = np.zeros(3, np.float32)
    got = orc.opacity(pos, idx, a, np.zeros(3, np.uint32), b, g)
    assert np.array_equal(got, np.array([_texel_alpha(2, 0), _texel_alpha(3, 0), _texel_alpha(0, 0)], np.float32))


def test_untextured_material_is_opaque():
    pos, idx, _ = _lookup_scene()
    a = orc.Alpha(np.zeros((3, 2)), None, {}, 1)
    assert orc.opacity(pos, idx, a, [0], [0.3], [0.3])[0] == 1.0


def _two_material_scene(n=3000, seed=7):
    sc = hiprt.HostScene.random(n, 0.03, seed)
    pos, idx, _ = sc.arrays()
    rng = np.random.default_rng(seed)
    mat = (rng.random(len(idx)) < 0.5).astype(np.uint32)
    tc = rng.uniform(-2.0, 3.0, (len(pos), 2)).astype(np.float32)
    sc = hiprt.HostScene.from_arrays(pos, idx, mat)
    return sc, pos, idx, mat, tc


def _texture(alpha, h=16, w=16):
    img = np.ones((h, w, 4), np.float32)
    img[..., 3] = alpha
    return img


def test_opaque_texture_changes_nothing():
    sc, pos, idx, mat, tc = _two_material_scene()
    bvh = hiprt.Bvh(sc)
    rays = uniform_sphere_rays(20_000, 3)
    a = orc.Alpha(tc, mat, {0: _texture(1.0), 1: _texture(0.75)}, 2)
    want = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    got = orc.trace_closest_alpha(pos, idx, bvh.nodes, bvh.root, bvh.index, rays, a)
    for x, y in zip(got, want):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(orc.trace_any_alpha(pos, idx, bvh.nodes, bvh.root, bvh.index, rays, a),
                          orc.trace_any(pos, idx, bvh.nodes, bvh.root, bvh.index, rays))


@pytest.mark.parametrize("alpha", [0.0, 0.5])
def test_transparent_material_equals_removing_it(alpha):
    sc, pos, idx, mat, tc = _two_material_scene()
    bvh = hiprt.Bvh(sc)
    rays = uniform_sphere_rays(20_000, 4)
    a = orc.Alpha(tc, mat, {1: _texture(alpha)}, 2)   # material 0 untextured, material 1 see-through
    got = orc.trace_closest_alpha(pos, idx, bvh.nodes, bvh.root, bvh.index, rays, a)
    keep = np.nonzero(mat == 0)[0]
    t, u, v, i = orc.seq_closest(pos, idx[keep], rays)
    hit = t < FLT_MAX
    assert hit.mean() > 0.2
    assert np.array_equal(got[0].view(np.uint32), t.view(np.uint32))
    ids = np.where(hit, keep[i], 0).astype(np.uint32)
    assert (got[3] != ids).mean() < 1e-4   # only equal-t ties may name another triangle
    occ = orc.trace_any_alpha(pos, idx, bvh.nodes, bvh.root, bvh.index, rays, a)
    assert np.array_equal(occ, orc.seq_any(pos, idx[keep], rays))
