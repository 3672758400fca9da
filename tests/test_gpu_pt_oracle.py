"""The GPU path tracer against the oracle's path loop (oracle.cpp:
orc_path_trace, pinned in tests/test_oracle_pt.py): per-pixel radiance and
rays per level on hall frames, same camera, same PCG32 streams.

Both sides trace bit-identical hits for identical rays (tests/test_gpu_trace.py)
and add one throughput x emission per path, so a pixel whose paths meet the
same triangles has the same bits on both sides.  Bounce directions go through
cosf/sinf, whose device (ocml) and host (glibc) results differ in the last
bit for some arguments; a path whose direction differs that way can meet
another triangle where it passes within about 1e-7 of an edge, or (for the
speculative and dynamic tracers) another triangle of an equal-t tie, and from
there its pixel differs.  The bar, per case:
  * at most MAX_DIFF_PATHS of the paths traced may have taken another way:
    the pixels whose bits differ are at most that fraction of all paths;
  * the image sums agree within SUM_REL, and the rays traced per level
    within COUNT_REL (a diverged path changes a level's count by at most one).
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc

pytestmark = pytest.mark.gpu

MAX_DIFF_PATHS = 2e-4
SUM_REL = 2e-3
COUNT_REL = 1e-4


@pytest.fixture(scope="module")
def hall(hip_ctx):
    sc = hiprt.HostScene.generate(3)
    pos, idx, mat = sc.arrays()
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
    yield dict(sc=sc, pos=pos, idx=idx, mat=mat, mats=sc.materials(), bvh=bvh, ds=ds, cam=sc.cam)
    ds.close()


def _oracle(h, cam, seed, n_frames, **kw):
    seeds = orc.frame_seeds(seed, 0, n_frames)
    b = h["bvh"]
    img, levels, _ = orc.path_trace(h["pos"], h["idx"], h["mat"], h["mats"], b.nodes, b.root, b.index, cam, seeds,
                                    **kw)
    return img, levels


CASES = [
    # name, PathTracer options, frame size, frames, render calls
    ("plain", dict(batched=False), (320, 180), 2, 2),
    ("batched, persistent-while-while, sorted queues, stack 4",
     dict(batched=True, tracer="persistent-while-while", stack_config=4), (320, 180), 2, 2),
    ("batched, persistent-while-while, shading order, one counter",
     dict(batched=True, tracer="persistent-while-while", sort_rays=False, work_chunks=1), (333, 97), 2, 2),
    ("batched, dynamic-while-while, sorted queues, stack 4, 64 samples in flight",
     dict(batched=True, stack_config=4, samples=64), (80, 45), 64, 1),
    ("batched, dynamic-while-while, 16 samples in flight, camera and first-bounce refill below 60 lanes",
     dict(batched=True, stack_config=4, samples=16, refill_coherent=60), (160, 90), 16, 1),
    ("batched, speculative, 4-wide tree, 8 samples in flight",
     dict(batched=True, tracer="persistent-speculative-while-while", samples=8, bvh4=True), (160, 90), 8, 1),
    ("plain, emitters reflect", dict(batched=False, emitters_reflect=True), (320, 180), 1, 1),
]


@pytest.mark.parametrize("name,opts,res,frames,calls", CASES, ids=[c[0] for c in CASES])
def test_path_tracer_matches_oracle(torch_cuda, hip_ctx, hall, name, opts, res, frames, calls):
    print(f"case: hall path tracer vs oracle, {name}, {res[0]}x{res[1]}, {frames} frames", flush=True)
    cam = hall["cam"].copy()
    cam[10], cam[11] = res
    opts = dict(opts)
    ds = hall["ds"]
    own = None
    if opts.pop("bvh4", False):
        own = ds = hiprt.DeviceScene(hip_ctx, hall["sc"], hall["bvh"], bvh_format=hiprt.BVH4)
    reflect = opts.get("emitters_reflect", False)
    pt = hiprt.PathTracer(ds, cam, seed=7, **opts)
    for f in range(calls):
        pt.render(f)
    got, got_levels = pt.image(), pt.ray_totals()
    pt.close()
    if own is not None:
        own.close()
    want, want_levels = _oracle(hall, cam, 7, frames, emitters_reflect=reflect)
    n_paths = res[0] * res[1] * frames
    differ = (got.view(np.uint32) != want.view(np.uint32)).any(-1)
    sum_rel = abs(float(got.sum(dtype=np.float64)) - float(want.sum(dtype=np.float64))) / float(
        want.sum(dtype=np.float64))
    level_diff = np.abs(got_levels.astype(np.int64) - want_levels.astype(np.int64))
    print({"pixels_differ": int(differ.sum()), "paths": n_paths, "sum_rel": sum_rel,
           "levels_gpu": got_levels.tolist(), "levels_oracle": want_levels.tolist()})
    assert np.isfinite(got).all() and want.sum() > 0
    assert want_levels[0] == n_paths and got_levels[0] == n_paths
    assert differ.sum() <= MAX_DIFF_PATHS * n_paths
    assert sum_rel <= SUM_REL
    assert (level_diff <= COUNT_REL * want_levels + 1).all()


def test_frame_tuner_choice_matches_oracle(torch_cuda, hip_ctx, hall):
    """hiprt_pt_autotune picks a configuration by timing trial frames; the
    image it renders meets the same bar against the oracle."""
    cam = hall["cam"].copy()
    cam[10], cam[11] = 160, 90
    ds4 = hiprt.DeviceScene(hip_ctx, hall["sc"], hall["bvh"], bvh_format=hiprt.BVH4)
    print("case: frame tuner on a 160x90 hall frame, up to 4 samples in flight", flush=True)
    kw, info = hiprt.pt_autotune(hall["ds"], ds4, cam, max_samples=4)
    print(kw, info)
    assert kw["samples"] in (1, 2, 4) and 1 <= kw["park_leaves"] <= 4 and info["candidates"] >= 12
    ds = ds4 if info["layout"] == "bvh4" else hall["ds"]
    pt = hiprt.PathTracer(ds, cam, seed=7, batched=True, **kw)
    pt.render(0)
    got, got_levels = pt.image(), pt.ray_totals()
    pt.close()
    ds4.close()
    want, want_levels = _oracle(hall, cam, 7, kw["samples"])
    n_paths = 160 * 90 * kw["samples"]
    differ = (got.view(np.uint32) != want.view(np.uint32)).any(-1)
    assert differ.sum() <= MAX_DIFF_PATHS * n_paths + 1
    assert np.abs(got_levels.astype(np.int64) - want_levels.astype(np.int64)).max() <= COUNT_REL * n_paths + 1


REGEN_CASES = [
    # name, PathTracer options, frame size, samples per render call
    ("one slot per pixel, sorted, persistent-while-while", dict(tracer="persistent-while-while", regen_pool=1),
     (160, 90), 8),
    ("three slots per pixel, sorted, dynamic-while-while, stack 4",
     dict(stack_config=4, regen_pool=3), (160, 90), 8),
    ("two slots per pixel, shading order, persistent-while-while",
     dict(tracer="persistent-while-while", regen_pool=2, sort_rays=False), (97, 61), 6),
    ("four slots per pixel, speculative, 4-wide tree", dict(tracer="persistent-speculative-while-while",
                                                            regen_pool=4, bvh4=True), (128, 72), 4),
]

# With more than one slot per pixel each slot sums its own samples and the
# slots are added at the end, so a pixel's sum is reassociated: per pixel the
# image agrees with the oracle within PIXEL_REL (S float additions of
# non-negative terms), except where a path took another way (MAX_DIFF_PATHS).
PIXEL_REL = 1e-5


@pytest.mark.parametrize("name,opts,res,samples", REGEN_CASES, ids=[c[0] for c in REGEN_CASES])
def test_regeneration_matches_oracle(torch_cuda, hip_ctx, hall, name, opts, res, samples):
    print(f"case: hall path regeneration vs oracle, {name}, {res[0]}x{res[1]}, {samples} samples", flush=True)
    cam = hall["cam"].copy()
    cam[10], cam[11] = res
    opts = dict(opts)
    ds = hall["ds"]
    own = None
    if opts.pop("bvh4", False):
        own = ds = hiprt.DeviceScene(hip_ctx, hall["sc"], hall["bvh"], bvh_format=hiprt.BVH4)
    pt = hiprt.PathTracer(ds, cam, seed=7, batched=True, samples=samples, **opts)
    pt.render(0)
    got, got_levels = pt.image(), pt.ray_totals()
    pt.close()
    if own is not None:
        own.close()
    want, want_levels = _oracle(hall, cam, 7, samples)
    n_paths = res[0] * res[1] * samples
    scale = np.maximum(np.abs(want).max(-1), 1e-30)
    far = np.abs(got - want).max(-1) > PIXEL_REL * scale
    level_diff = np.abs(got_levels.astype(np.int64) - want_levels.astype(np.int64))
    print({"pixels_beyond_rel": int(far.sum()), "paths": n_paths, "levels_gpu": got_levels.tolist(),
           "levels_oracle": want_levels.tolist()})
    assert np.isfinite(got).all() and want.sum() > 0
    assert want_levels[0] == n_paths and got_levels[0] == n_paths
    assert far.sum() <= MAX_DIFF_PATHS * n_paths
    assert (level_diff <= COUNT_REL * want_levels + 1).all()


def test_regeneration_with_one_slot_equals_samples_in_flight(torch_cuda, hip_ctx, hall):
    """One slot per pixel adds a pixel's samples in sample order, as the
    in-flight schedule does: with the deterministic tracer the two images
    and their per-level ray counts are identical, over two render calls."""
    cam = hall["cam"].copy()
    cam[10], cam[11] = 200, 120
    print("case: regeneration (1 slot per pixel) vs 16 samples in flight, 200x120, 2 render calls", flush=True)
    out = []
    for kw in (dict(samples=16), dict(samples=16, regen_pool=1)):
        pt = hiprt.PathTracer(hall["ds"], cam, seed=11, batched=True, tracer="persistent-while-while",
                              stack_config=4, **kw)
        pt.render(0)
        pt.render(1)
        out.append((pt.image(), pt.ray_totals(), pt.ray_counts()))
        pt.close()
    (a, ta, ca), (b, tb, cb) = out
    assert a.sum() > 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ta, tb) and np.array_equal(ca, cb)


def test_path_tracer_with_shading_normals_matches_oracle(torch_cuda, hip_ctx, hall):
    """The hall with random unit vertex normals: every bounce flips by the
    interpolated shading normal (flip_normals_to_ray), on the GPU (plain and
    batched schedules) as in the oracle's path loop; the image differs from
    the one without normals."""
    sc = hiprt.HostScene.generate(3)
    pos, idx, mat = sc.arrays()
    rng = np.random.default_rng(8)
    nrm = rng.normal(size=(len(pos), 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    sc.set_normals(nrm)
    ds = hiprt.DeviceScene(hip_ctx, sc, hall["bvh"])
    cam = hall["cam"].copy()
    cam[10], cam[11] = 160, 90
    want, want_levels = _oracle(hall, cam, 5, 2, normals=nrm)
    plain_img, _ = _oracle(hall, cam, 5, 2)
    n_paths = 160 * 90 * 2
    for opts in (dict(batched=False), dict(batched=True, tracer="persistent-while-while", samples=2)):
        print(f"case: hall 160x90 with random vertex normals, {opts}", flush=True)
        pt = hiprt.PathTracer(ds, cam, seed=5, **opts)
        for f in range(2 // opts.get("samples", 1)):
            pt.render(f)
        got, got_levels = pt.image(), pt.ray_totals()
        pt.close()
        differ = (got.view(np.uint32) != want.view(np.uint32)).any(-1)
        print({"pixels_differ": int(differ.sum()), "levels_gpu": got_levels.tolist(),
               "levels_oracle": want_levels.tolist()})
        assert differ.sum() <= MAX_DIFF_PATHS * n_paths + 1
        assert (np.abs(got_levels.astype(np.int64) - want_levels.astype(np.int64)) <= COUNT_REL * want_levels + 1).all()
    ds.close()
    # the normals change most lit pixels
    lit = (want.sum(-1) > 0) | (plain_img.sum(-1) > 0)
    assert (plain_img.view(np.uint32) != want.view(np.uint32)).any(-1)[lit].mean() > 0.3
