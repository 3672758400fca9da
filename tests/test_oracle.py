"""The CPU oracle pinned against closed forms, special cases and brute force.

Each check is chosen so that a plausible restatement mistake -- a dropped
term, a flipped comparison, an inclusive bound where the reference has an
exclusive one, the wrong child first on a distance tie -- fails it.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, uniform_sphere_rays


def one_ray(o, d, tmin=1e-4, tmax=float(FLT_MAX)):
    r = np.zeros((8, 1), np.float32)
    r[0:3, 0] = o
    r[3, 0] = tmin
    r[4:7, 0] = d
    r[7, 0] = tmax
    return r


def tri_scene(verts):
    pos = np.asarray(verts, np.float32).reshape(-1, 3)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    return pos, idx


# ---------------------------------------------------------------- triangle test

def test_centroid_hit_closed_form():
    # triangle in the plane z = 2; ray from the origin through the centroid
    pos, idx = tri_scene([[0, 0, 2], [1, 0, 2], [0, 1, 2]])
    c = np.array([1 / 3, 1 / 3, 2], np.float64)
    d = (c / np.linalg.norm(c)).astype(np.float32)
    t, u, v, i = orc.seq_closest(pos, idx, one_ray([0, 0, 0], d))
    assert i[0] == 0 and t[0] != FLT_MAX
    assert abs(t[0] - np.linalg.norm(c)) < 1e-6 * np.linalg.norm(c)
    assert abs(u[0] - 1 / 3) < 1e-6 and abs(v[0] - 1 / 3) < 1e-6


def test_barycentrics_name_vertices_b_and_c():
    # hit point close to vertex B must give beta ~ 1, close to C gamma ~ 1
    pos, idx = tri_scene([[0, 0, 1], [4, 0, 1], [0, 4, 1]])
    t, u, v, _ = orc.seq_closest(pos, idx, one_ray([3.5, 0.25, 0], [0, 0, 1]))
    assert abs(u[0] - 0.875) < 1e-6 and abs(v[0] - 0.0625) < 1e-6 and abs(t[0] - 1) < 1e-6
    t, u, v, _ = orc.seq_closest(pos, idx, one_ray([0.25, 3.5, 0], [0, 0, 1]))
    assert abs(u[0] - 0.0625) < 1e-6 and abs(v[0] - 0.875) < 1e-6


def test_edge_rule_is_asymmetric():
    # beta > 0 and gamma > 0 are strict, beta + gamma <= 1 is inclusive
    # (libgi/intersect.h:61).  Axis-aligned unit triangle, rays along -z.
    pos, idx = tri_scene([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    on_ab = one_ray([0.5, 0.0, 1], [0, 0, -1])     # gamma == 0 -> miss
    on_ac = one_ray([0.0, 0.5, 1], [0, 0, -1])     # beta == 0 -> miss
    on_bc = one_ray([0.5, 0.5, 1], [0, 0, -1])     # beta + gamma == 1 -> hit
    assert orc.seq_closest(pos, idx, on_ab)[0][0] == FLT_MAX
    assert orc.seq_closest(pos, idx, on_ac)[0][0] == FLT_MAX
    t, u, v, _ = orc.seq_closest(pos, idx, on_bc)
    assert t[0] == 1.0 and u[0] == 0.5 and v[0] == 0.5


def test_t_interval_is_open_on_both_ends():
    pos, idx = tri_scene([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    o, d = [0.25, 0.25, 1], [0, 0, -1]
    assert orc.seq_closest(pos, idx, one_ray(o, d, tmin=1.0))[0][0] == FLT_MAX
    assert orc.seq_closest(pos, idx, one_ray(o, d, tmax=1.0))[0][0] == FLT_MAX
    assert orc.seq_closest(pos, idx, one_ray(o, d, tmin=0.999, tmax=1.001))[0][0] == 1.0
    # behind the origin: t = -1 is never accepted
    assert orc.seq_closest(pos, idx, one_ray([0.25, 0.25, -1], d))[0][0] == FLT_MAX


def test_degenerate_and_edge_on_never_hit():
    # collinear vertices: M = 0
    pos, idx = tri_scene([[0, 0, 0], [1, 1, 1], [2, 2, 2]])
    assert orc.seq_closest(pos, idx, one_ray([1, 1, 2], [0, 0, -1]))[0][0] == FLT_MAX
    # ray inside the triangle's plane: M = 0
    pos, idx = tri_scene([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert orc.seq_closest(pos, idx, one_ray([-1, 0.25, 0], [1, 0, 0]))[0][0] == FLT_MAX


def test_closest_of_two_and_strict_tie_keeps_first():
    # two parallel triangles; the nearer wins regardless of order
    pos, idx = tri_scene([[0, 0, 3], [1, 0, 3], [0, 1, 3], [0, 0, 2], [1, 0, 2], [0, 1, 2]])
    t, _, _, i = orc.seq_closest(pos, idx, one_ray([0.2, 0.2, 0], [0, 0, 1]))
    assert i[0] == 1 and t[0] == 2.0
    # the same triangle twice: strict "<" keeps the first index
    pos, idx = tri_scene([[0, 0, 2], [1, 0, 2], [0, 1, 2]] * 2)
    assert orc.seq_closest(pos, idx, one_ray([0.2, 0.2, 0], [0, 0, 1]))[3][0] == 0


# ---------------------------------------------------------------- slab test

def _bvh_one_box(box_l, box_r):
    """A root with two leaf children holding triangles 0 and 1."""
    nodes = np.zeros(3, hiprt.BBVH_NODE)
    nodes[0]["box_l"] = box_l
    nodes[0]["box_r"] = box_r
    nodes[0]["link_l"], nodes[0]["link_r"] = 1, 2
    nodes[1]["link_l"], nodes[1]["link_r"] = 0, -1
    nodes[2]["link_l"], nodes[2]["link_r"] = -1, -1
    return nodes, np.array([0, 1], np.uint32)


def test_box_test_culls_and_orders():
    # Two unit squares at z=1 (left box) and z=2 (right box).  A ray hitting
    # both returns the nearer one; a box entirely behind t_max is culled.
    pos, idx = tri_scene([[0, 0, 1], [1, 0, 1], [0, 1, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2]])
    nodes, index = _bvh_one_box([[0, 0, 1], [1, 1, 1]], [[0, 0, 2], [1, 1, 2]])
    t, _, _, i = orc.trace_closest(pos, idx, nodes, 0, index, one_ray([0.2, 0.2, 0], [0, 0, 1]))
    assert i[0] == 0 and t[0] == 1.0
    # t_max 1.5: only the first box (and triangle) is reachable
    t, _, _, i = orc.trace_closest(pos, idx, nodes, 0, index, one_ray([0.2, 0.2, 0], [0, 0, 1], tmax=1.5))
    assert i[0] == 0
    t, _, _, i = orc.trace_closest(pos, idx, nodes, 0, index, one_ray([0.2, 0.2, 0], [0, 0, 1], tmax=0.5))
    assert t[0] == FLT_MAX
    # boxes in the wrong place hide the triangles: traversal must trust the boxes
    nodes2, _ = _bvh_one_box([[5, 5, 5], [6, 6, 6]], [[5, 5, 5], [6, 6, 6]])
    assert orc.trace_closest(pos, idx, nodes2, 0, index, one_ray([0.2, 0.2, 0], [0, 0, 1]))[0][0] == FLT_MAX


def test_axis_parallel_ray_uses_infinite_inverse():
    # d.x = d.y = 0: id = +-inf; origin inside the x/y slabs -> hit, outside -> miss
    pos, idx = tri_scene([[0, 0, 1], [1, 0, 1], [0, 1, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1]])
    nodes, index = _bvh_one_box([[0, 0, 1], [1, 1, 1]], [[0, 0, 1], [1, 1, 1]])
    assert orc.trace_closest(pos, idx, nodes, 0, index, one_ray([0.2, 0.2, 0], [0, 0, 1]))[0][0] == 1.0
    assert orc.trace_closest(pos, idx, nodes, 0, index, one_ray([1.5, 0.2, 0], [0, 0, 1]))[0][0] == FLT_MAX
    # origin exactly on the slab plane x = 0 with d.x = 0: (0 * inf) = NaN; the
    # select chain of intersect4 then keeps the box (NaN compares false), and
    # the triangle test rejects beta == 0, so no hit -- but no crash either
    t = orc.trace_closest(pos, idx, nodes, 0, index, one_ray([0.0, 0.2, 0], [0, 0, 1]))[0][0]
    assert t == FLT_MAX


def test_origin_inside_box_enters_it():
    pos, idx = tri_scene([[-1, -1, 0.5], [1, -1, 0.5], [-1, 1, 0.5], [-1, -1, 0.5], [1, -1, 0.5], [-1, 1, 0.5]])
    nodes, index = _bvh_one_box([[-1, -1, -1], [1, 1, 1]], [[-1, -1, -1], [1, 1, 1]])
    t, _, _, _ = orc.trace_closest(pos, idx, nodes, 0, index, one_ray([-0.5, -0.5, 0], [0, 0, 1]))
    assert t[0] == 0.5


# ---------------------------------------------------------------- traversal vs brute force

@pytest.mark.parametrize("split", [hiprt.SPLIT_OM, hiprt.SPLIT_SM, hiprt.SPLIT_SAH])
@pytest.mark.parametrize("esc", [False, True])
@pytest.mark.parametrize("max_tris", [1, 4])
def test_bvh_closest_equals_brute_force(split, esc, max_tris):
    sc = hiprt.HostScene.random(1500, sigma=0.03, seed=7 + split)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc, split=split, max_tris=max_tris, planes=16, esc=esc)
    rays = uniform_sphere_rays(4000, seed=split * 10 + max_tris, lo=-0.2, hi=1.2)
    a = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    b = orc.seq_closest(pos, idx, rays)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))   # same t, always
    diff = a[3] != b[3]
    assert diff.sum() == 0 or np.all(a[0][diff] == b[0][diff])          # ids differ only at equal t
    occ = orc.trace_any(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    assert np.array_equal(occ, orc.seq_any(pos, idx, rays))
    assert np.array_equal(occ.astype(bool), a[0] != FLT_MAX)


def test_counts_match_visits():
    sc = hiprt.HostScene.random(500, seed=3)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    rays = uniform_sphere_rays(2000, seed=5)
    *_, counts = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays, counts=True)
    inner = int((bvh.nodes["link_r"] >= 0).sum())
    assert counts[:, 0].max() <= inner and counts[:, 1].max() <= len(bvh.index)
    assert counts[:, 0].min() >= 1   # the root is always visited


# ---------------------------------------------------------------- FP64 oracle

def test_fp64_agrees_with_fp32_hits():
    sc = hiprt.HostScene.random(2000, seed=9)
    pos, idx, _ = sc.arrays()
    rays = uniform_sphere_rays(5000, seed=9)
    t, u, v, i = orc.seq_closest(pos, idx, rays)
    ok = t != FLT_MAX
    t64, b64, g64, inside = orc.fp64_intersect(pos, idx, i[ok], rays[:, ok])
    assert inside.all()
    assert np.median(np.abs(t[ok] - t64) / t64) < 2e-7
    assert np.max(np.abs(u[ok] - b64)) < 1e-4 and np.max(np.abs(v[ok] - g64)) < 1e-4


def test_fp64_closed_form():
    pos, idx = tri_scene([[0, 0, 2], [1, 0, 2], [0, 1, 2]])
    t64, b64, g64, inside = orc.fp64_intersect(pos, idx, np.array([0], np.uint32), one_ray([0.25, 0.5, 0], [0, 0, 1]))
    assert t64[0] == 2.0 and b64[0] == 0.25 and g64[0] == 0.5 and inside[0] == 1


# ---------------------------------------------------------------- ray generation

@pytest.mark.parametrize("jitter", [False, True])
def test_camera_rays_closed_form(jitter):
    w, h, fovy = 64, 48, 60.0
    rng = np.random.default_rng(3)
    ox = rng.uniform(-0.5, 0.5, w * h).astype(np.float32) if jitter else np.zeros(w * h, np.float32)
    oy = rng.uniform(-0.5, 0.5, w * h).astype(np.float32) if jitter else np.zeros(w * h, np.float32)
    rays = orc.camera_rays([0.5, 0.5, -1], [0, 0, 1], [0, 1, 0], fovy, w, h, *((ox, oy) if jitter else ()))
    d = rays[4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=0), 1.0, atol=1e-6)
    assert np.all(rays[0] == np.float32(0.5)) and np.all(rays[2] == np.float32(-1))
    assert np.all(rays[3] == np.float32(1e-4)) and np.all(rays[7] == FLT_MAX)
    # pinhole: d ~ dir + U u + V v with U = dir x up = (-1,0,0), V = (0,1,0)
    nh = np.tan(np.pi * fovy / 360.0)
    nw = w / h * nh
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    u = ((xs + 0.5 + ox.reshape(h, w)) / w * 2 - 1) * nw
    v = ((ys + 0.5 + oy.reshape(h, w)) / h * 2 - 1) * nh
    q = np.stack([-u, v, np.ones_like(u)]).reshape(3, -1)
    q /= np.linalg.norm(q, axis=0)
    assert np.max(np.abs(q - d)) < 2e-6
    # fovy is vertical and in degrees: d.y / d.z of pixel row y is
    # ((y + 0.5 + oy) / h * 2 - 1) * tan(fovy / 2), independent of the column
    ratio = (d[1] / d[2]).reshape(h, w)
    want = ((np.arange(h)[:, None] + 0.5 + oy.reshape(h, w)) / h * 2 - 1) * np.tan(np.radians(fovy / 2))
    assert np.max(np.abs(ratio - want)) < 1e-6


def test_product_camera_rays_equal_oracle():
    cam = hiprt.HostScene.generate(1).cam.copy()
    cam[10], cam[11] = 160, 120
    a = hiprt.camera_rays(cam)
    b = orc.camera_rays(cam[0:3], cam[3:6], cam[6:9], float(cam[9]), 160, 120)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_cosine_rays_distribution():
    # one big triangle in z = 0 hit from above; bounce directions must lie in
    # the upper hemisphere, be unit length, and have cos^2(theta) ~ U[0,1]
    n = 200_000
    pos, idx = tri_scene([[-100, -100, 0], [100, -100, 0], [-100, 100, 0]])
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = [[-1], [-1], [1]]
    rays[3] = 1e-4
    rays[4:7] = [[0], [0], [-1]]
    rays[7] = FLT_MAX
    t, u, v, i = orc.seq_closest(pos, idx, rays[:, :1])
    rng = np.random.default_rng(0)
    out = orc.cosine_rays(pos, idx, rays, np.repeat(t, n), np.repeat(u, n), np.repeat(v, n), np.repeat(i, n),
                          rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32))
    d = out[4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=0), 1.0, atol=1e-5)
    assert np.all(d[2] >= 0)
    c2 = np.sort(d[2] ** 2)
    ks = np.max(np.abs(c2 - np.arange(1, n + 1) / n))
    assert ks < 0.01
    # hit point is the origin of the bounce ray
    assert np.allclose(out[0:3, 0], [-1, -1, 0], atol=1e-6)
    # a ray arriving from below flips the normal: bounce into z < 0
    rays[4:7] = [[0], [0], [1]]
    rays[0:3] = [[-1], [-1], [-1]]
    t, u, v, i = orc.seq_closest(pos, idx, rays[:, :1])
    out = orc.cosine_rays(pos, idx, rays[:, :100], np.repeat(t, 100), np.repeat(u, 100), np.repeat(v, 100),
                          np.repeat(i, 100), rng.random(100, dtype=np.float32), rng.random(100, dtype=np.float32))
    assert np.all(out[6] <= 0)
