"""Pins of the oracle's path loop (oracle.cpp: orc_path_trace) and of PCG32.

* PCG32: the published known answer of the PCG family's minimal C
  implementation (pcg32_srandom_r(42, 54) -> 0xa15c02b7 0x7b47f409 ...), and
  the numpy restatement of the kernels' streams (tests/pcg.py);
* closed forms on a unit box: a floor of albedo a under five emitting faces
  seen from inside gives exactly a * e in every pixel (the camera ray hits
  the floor, the cosine bounce leaves upward and ends at an emitter: the path
  stops there as gi/pt.cpp:86-89 has it); with emitters reflecting the paths
  go on and add more; a black floor ends every path after the first hit;
* Russian roulette: in a closed box that never emits, the rays per level are
  exactly the pixels whose PCG32 streams keep them (third draw < 0.8 at
  levels 2 and 3), counted independently with numpy;
* composition: on a random scene with emitters, the image and per-level ray
  counts equal a numpy loop over the separately pinned steps (camera rays,
  BVH closest hit, cosine bounce rays, PCG32 streams), bit for bit.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from pcg import Pcg32, stream_id

FLT_MAX = np.float32(3.4028234663852886e38)


def test_pcg32_known_answer():
    got = orc.pcg_sequence(42, 54, 6)
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert got.tolist() == want


def test_pcg32_streams_equal_numpy_restatement():
    ids = stream_id(123, np.arange(64, dtype=np.uint64), 3)
    g = Pcg32(ids)
    want = np.stack([g.next_float() for _ in range(5)], 1)
    for k in (0, 17, 63):
        got = orc.pcg_sequence(0x853C49E6748FEA9B, int(ids[k]), 5, raw=False)
        assert np.array_equal(got.view(np.uint32), want[k].view(np.uint32))


def _box(floor_mat, other_mat):
    """Unit cube, two triangles per face; the floor (y = 0) gets floor_mat."""
    P = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)
    v = lambda x, y, z: 4 * x + 2 * y + z
    quads = {
        "floor": (v(0, 0, 0), v(1, 0, 0), v(1, 0, 1), v(0, 0, 1)),
        "ceil": (v(0, 1, 0), v(0, 1, 1), v(1, 1, 1), v(1, 1, 0)),
        "x0": (v(0, 0, 0), v(0, 0, 1), v(0, 1, 1), v(0, 1, 0)),
        "x1": (v(1, 0, 0), v(1, 1, 0), v(1, 1, 1), v(1, 0, 1)),
        "z0": (v(0, 0, 0), v(0, 1, 0), v(1, 1, 0), v(1, 0, 0)),
        "z1": (v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)),
    }
    idx, mat = [], []
    for name, (a, b, c, d) in quads.items():
        idx += [(a, b, c), (a, c, d)]
        mat += [floor_mat if name == "floor" else other_mat] * 2
    return P, np.array(idx, np.uint32), np.array(mat, np.uint32)


def _tree(pos, idx, mat=None):
    sc = hiprt.HostScene.from_arrays(pos, idx, mat)
    return sc, hiprt.Bvh(sc)


# camera inside the box, looking down at the floor, off-centre so that no
# pixel's ray meets the floor's diagonal exactly
_BOX_CAM = np.array([0.41, 0.8, 0.52, 0, -1, 0, 0, 0, 1, 50, 32, 32], np.float32)


def _run(pos, idx, mat, mats, seeds, cam=_BOX_CAM, **kw):
    sc, bvh = _tree(pos, idx, mat)
    return orc.path_trace(pos, idx, mat, mats, bvh.nodes, bvh.root, bvh.index, cam, seeds, **kw)


def test_floor_under_emitters_is_albedo_times_emission():
    pos, idx, mat = _box(0, 1)
    mats = np.array([[0.5, 0.5, 0.5, 0, 0, 0, 0, 0], [0.3, 0.3, 0.3, 0, 2, 3, 4, 0]], np.float32)
    seeds = orc.frame_seeds(3, 0, 2)
    img, levels, _ = _run(pos, idx, mat, mats, seeds)
    n = 32 * 32
    # two frames: a * e added twice
    assert np.array_equal(img.reshape(-1, 3), np.tile(np.float32([2.0, 3.0, 4.0]), (n, 1)))
    assert levels.tolist() == [2 * n, 2 * n, 0, 0, 0]


def test_emitters_reflect_is_an_option():
    pos, idx, mat = _box(0, 1)
    mats = np.array([[0.5, 0.5, 0.5, 0, 0, 0, 0, 0], [0.3, 0.3, 0.3, 0, 2, 3, 4, 0]], np.float32)
    seeds = orc.frame_seeds(3, 0, 1)
    img, levels, _ = _run(pos, idx, mat, mats, seeds, emitters_reflect=True)
    n = 32 * 32
    assert levels[1] == n and levels[2] > 0.9 * n
    assert (img.reshape(-1, 3) >= np.float32([1.0, 1.5, 2.0])).all()
    assert img.sum() > 1.2 * n * 4.5


def test_floor_normals_decide_the_flip():
    """flip_normals_to_ray flips by the shading normal (libgi/util.h:226-232).
    The box's floor is wound with its geometric normal pointing down (out of
    the box).  With vertex normals pointing up, a camera ray from above is not
    on their side, so nothing flips, the cosine bounce leaves about the
    downward geometric normal, out of the box, and no path meets an emitter.
    With the normals down (equal to the geometric normal) both flip, the
    bounce goes up and every pixel is a * e, as without normals.  Flipping by
    the geometric normal would light the floor in both."""
    pos, idx, mat = _box(0, 1)
    mats = np.array([[0.5, 0.5, 0.5, 0, 0, 0, 0, 0], [0.3, 0.3, 0.3, 0, 2, 3, 4, 0]], np.float32)
    seeds = orc.frame_seeds(3, 0, 1)
    n = 32 * 32
    up = np.tile(np.float32([[0, 1, 0]]), (len(pos), 1))
    img, levels, _ = _run(pos, idx, mat, mats, seeds, normals=up)
    assert levels.tolist() == [n, n, 0, 0, 0] and not img.any()
    img, levels, _ = _run(pos, idx, mat, mats, seeds, normals=-up)
    assert np.array_equal(img.reshape(-1, 3), np.tile(np.float32([1.0, 1.5, 2.0]), (n, 1)))


def test_black_floor_ends_every_path():
    pos, idx, mat = _box(0, 1)
    mats = np.array([[0, 0, 0, 0, 0, 0, 0, 0], [0.3, 0.3, 0.3, 0, 2, 3, 4, 0]], np.float32)
    img, levels, _ = _run(pos, idx, mat, mats, orc.frame_seeds(1, 0, 1))
    assert levels.tolist() == [32 * 32, 0, 0, 0, 0]
    assert not img.any()


def test_roulette_counts_follow_the_streams():
    pos, idx, mat = _box(0, 0)
    mats = np.array([[0.5, 0.5, 0.5, 0, 0, 0, 0, 0]], np.float32)
    seeds = orc.frame_seeds(9, 4, 1)
    img, levels, _ = _run(pos, idx, mat, mats, seeds)
    n = 32 * 32
    px = np.arange(n, dtype=np.uint64)

    def survive(level):
        g = Pcg32(stream_id(int(seeds[0]), px, level))
        g.next_float(), g.next_float()
        return g.next_float() < np.float32(0.8)

    s2, s3 = survive(2), survive(3)
    assert levels.tolist() == [n, n, n, int(s2.sum()), int((s2 & s3).sum())]
    assert not img.any()


def _composed(pos, idx, mat, mats, bvh, cam, seeds, max_depth=4, rr_start=3, p=0.8, jitter=True):
    """The path loop as a numpy loop over the oracle's separately pinned steps."""
    w, h = int(cam[10]), int(cam[11])
    n = w * h
    img = np.zeros((n, 3), np.float32)
    counts = np.zeros(max_depth + 1, np.uint64)
    scale_rr = np.float32(1.0) / np.float32(p)
    for seed in seeds:
        alive = np.arange(n, dtype=np.uint64)
        off = ()
        if jitter:   # sub-pixel offsets: first two uniforms of stream (seed, pixel, 255)
            g = Pcg32(stream_id(int(seed), alive, 255))
            off = (g.next_float() - np.float32(0.5), g.next_float() - np.float32(0.5))
        rays = orc.camera_rays(cam[0:3], cam[3:6], cam[6:9], float(cam[9]), w, h, *off)
        T = np.ones((n, 3), np.float32)
        for L in range(max_depth + 1):
            counts[L] += len(alive)
            t, u, v, tri = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
            hit = t != FLT_MAX
            m = mats[mat[np.where(hit, tri, 0)]]
            emi = m[:, 4:7]
            em = hit & (emi != 0).any(1)
            img[alive[em]] += T[em] * emi[em]
            cont = hit & ~em
            if L >= max_depth:
                break
            g = Pcg32(stream_id(int(seed), alive, L))
            xi0, xi1 = g.next_float(), g.next_float()
            scale = np.float32(1.0)
            if L + 1 >= rr_start:
                cont &= g.next_float() < np.float32(p)
                scale = scale_rr
            nxt = orc.cosine_rays(pos, idx, rays, t, u, v, tri, xi0, xi1)
            T = (T * m[:, 0:3]) * scale
            cont &= (T != 0).any(1)
            rays, alive, T = np.ascontiguousarray(nxt[:, cont]), alive[cont], T[cont]
            if not len(alive):
                break
    return img.reshape(h, w, 3), counts


@pytest.mark.parametrize("max_depth,rr_start,jitter", [(4, 3, True), (2, 1, True), (4, 3, False)])
def test_path_loop_equals_composed_steps(max_depth, rr_start, jitter):
    sc = hiprt.HostScene.random(3000, 0.03, seed=21)
    pos, idx, _ = sc.arrays()
    k = np.arange(len(idx))
    mat = np.where(k % 7 == 0, 1, np.where(k % 3 == 1, 2, 0)).astype(np.uint32)
    mats = np.array([[0.7, 0.6, 0.5, 0, 0, 0, 0, 0], [0.2, 0.2, 0.2, 0, 5, 4, 3, 0],
                     [0.9, 0.9, 0.9, 0, 0, 0, 0, 0]], np.float32)
    _, bvh = _tree(pos, idx, mat)
    cam = np.array([0.5, 0.5, -1, 0, 0, 1, 0, 1, 0, 60, 48, 36], np.float32)
    seeds = orc.frame_seeds(5, 0, 3)
    img, levels, _ = orc.path_trace(pos, idx, mat, mats, bvh.nodes, bvh.root, bvh.index, cam, seeds,
                                    max_depth=max_depth, rr_start=rr_start, jitter=jitter)
    want_img, want_levels = _composed(pos, idx, mat, mats, bvh, cam, seeds, max_depth, rr_start, jitter=jitter)
    assert levels.tolist() == want_levels.tolist()
    assert levels[1] > 0 and img.sum() > 0
    assert np.array_equal(img.view(np.uint32), want_img.view(np.uint32))
