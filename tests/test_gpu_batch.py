"""4-wide BVH traversal, the multi-scene descriptor-table launch and the
autotuner, all held to the oracle: identical hits except between triangles
with exactly equal t (the traversal order differs from the reference's)."""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
import workloads
from parity import FLT_MAX, classify, uniform_sphere_rays

pytestmark = pytest.mark.gpu


def _np_hits(h):
    return (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(), h["id"].cpu().numpy().view(np.uint32))


def _same_up_to_ties(got, want):
    bits = lambda a: np.ascontiguousarray(a).view(np.uint32)
    diff = got[3] != want[3]
    assert np.array_equal(bits(got[0]), bits(want[0]))
    assert int(diff.sum()) == int((diff & (bits(got[0]) == bits(want[0]))).sum())
    same = ~diff & (got[0] != FLT_MAX)
    for a, b in zip(got[1:3], want[1:3]):
        assert np.array_equal(bits(a)[same], bits(b)[same])


@pytest.mark.parametrize("cfg", [1, 3])
@pytest.mark.parametrize("tracer", ["persistent-while-while", "dynamic-while-while"])
def test_bvh4_matches_oracle(torch_cuda, hip_ctx, cfg, tracer):
    print(f"case: BVH4 config {cfg}, {tracer}", flush=True)
    torch = torch_cuda
    sc = hiprt.HostScene.generate(cfg)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh, hiprt.BVH4, hiprt.TRI_EXACT)
    cam = sc.cam.copy()
    cam[10], cam[11] = 320, 240
    rays = np.concatenate([hiprt.camera_rays(cam), uniform_sphere_rays(50_000, 3, lo=0.0,
                                                                        hi=1.0 if cfg == 1 else 15.0)], 1)
    h = hiprt.new_hits(rays.shape[1])
    hiprt.trace_closest(ds, torch.from_numpy(rays).cuda(), h, tracer=tracer)
    torch.cuda.synchronize()
    want = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    _same_up_to_ties(_np_hits(h), want)
    occ = torch.empty(rays.shape[1], dtype=torch.uint8, device="cuda")
    short = rays.copy()
    short[7] = np.float32(0.5)
    hiprt.trace_any(ds, torch.from_numpy(short).cuda(), occ, tracer=tracer)
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy(), orc.trace_any(pos, idx, bvh.nodes, bvh.root, bvh.index, short))
    ds.close()


def test_plain_kernel_refuses_bvh4(torch_cuda, hip_ctx):
    sc = hiprt.HostScene.generate(1)
    ds = hiprt.DeviceScene(hip_ctx, sc, hiprt.Bvh(sc), hiprt.BVH4)
    rays = torch_cuda.from_numpy(uniform_sphere_rays(100, 1)).cuda()
    with pytest.raises(hiprt.HiprtError, match="persistent"):
        hiprt.trace_closest(ds, rays, hiprt.new_hits(100), tracer="while-while")


@pytest.mark.parametrize("layout", [hiprt.BVH2, hiprt.BVH4])
def test_batch_of_uneven_scenes_one_launch(torch_cuda, hip_ctx, layout):
    print(f"case: 12 uneven random scenes, one batched launch, layout {layout}", flush=True)
    torch = torch_cuda
    work = workloads.multi_scene(12, seed=4, tri_hi=60_000, ray_hi=150_000)
    entries, refs = [], []
    for sc, bvh, rays in work:
        ds = hiprt.DeviceScene(hip_ctx, sc, bvh, layout)
        d_rays = torch.from_numpy(rays).cuda()
        h = hiprt.new_hits(rays.shape[1])
        for x in h.values():
            x.fill_(-3)
        entries.append((ds, d_rays, h, None))
        pos, idx, _ = sc.arrays()
        refs.append(orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays))
    hiprt.trace_batch(entries, tracer="dynamic-while-while")
    torch.cuda.synchronize()
    for (ds, _, h, _), want in zip(entries, refs):
        _same_up_to_ties(_np_hits(h), want)
    # any hit through the same table
    occs = []
    for k, (ds, d_rays, h, _) in enumerate(entries):
        occs.append(torch.full((d_rays.shape[1],), 9, dtype=torch.uint8, device="cuda"))
    hiprt.trace_batch([(e[0], e[1], None, o) for e, o in zip(entries, occs)], any=True,
                      tracer="persistent-speculative-while-while")
    torch.cuda.synchronize()
    for o, want in zip(occs, refs):
        assert np.array_equal(o.cpu().numpy().astype(bool), want[0] != FLT_MAX)


def test_autotuner_picks_a_config_with_identical_results(torch_cuda, hip_ctx):
    print("case: autotune the hall on 131,072 diffuse rays", flush=True)
    torch = torch_cuda
    sc = hiprt.HostScene.generate(3)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
    cam = torch.empty((8, 1920 * 1080), dtype=torch.float32, device="cuda")
    hiprt.camera_rays_device(sc.cam, cam)
    h = hiprt.new_hits(cam.shape[1])
    hiprt.trace_closest(ds, cam, h, tracer="while-while")
    diff = torch.empty_like(cam)
    hiprt.bounce_cosine(ds, cam, h, diff, seed=3, level=0)
    sample = diff[:, ::16].contiguous()[:, :131072].contiguous()
    res, tuned = hiprt.autotune(hip_ctx, sc, bvh, sample)
    print(res)
    assert res["candidates"] >= 20 and res["mrays_per_s"] > 0
    h2 = hiprt.new_hits(sample.shape[1])
    hiprt.trace_closest(tuned, sample, h2, tracer=res["tracer"], refill_threshold=res["refill_threshold"],
                        blocks=res["blocks"], stack_config=res["stack_config"])
    torch.cuda.synchronize()
    want = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, sample.cpu().numpy())
    _same_up_to_ties(_np_hits(h2), want)
    tuned.close()
