"""The direct-lighting steps on the GPU (hiprt_sample_direct,
hiprt_integrate_dir_sample, hiprt_integrate_light_sample) against the
oracle's wavefront direct lighting (oracle.cpp: orc_direct, pinned in
tests/test_oracle_direct.py), on hall frames with their emissive arches.

Light mode uses only correctly rounded operations (the light pick, the point
on the triangle, normalisation, the any-hit shadow test), so its framebuffer
equals the oracle's bit for bit.  The uniform and cosine modes turn their
samples through cosf/sinf, whose device and host results may differ in the
last bit: the cosine weight (and the cosine pdf) of such a sample then
differs by about an ulp, so each pixel must agree within PIXEL_REL; and a
direction that differs that way can meet another triangle, which at most
MAX_DIFF of the samples may do (pixels outside PIXEL_REL).  The framebuffer
sums agree within SUM_REL."""
import numpy as np
import pytest

import hiprt
import oracle_py as orc

pytestmark = pytest.mark.gpu

MAX_DIFF = 2e-4
PIXEL_REL = 1e-5
SUM_REL = 1e-3


@pytest.fixture(scope="module", params=[2, 3])
def hall(request, hip_ctx):
    sc = hiprt.HostScene.generate(request.param)
    pos, idx, mat = sc.arrays()
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
    yield dict(cfg=request.param, sc=sc, pos=pos, idx=idx, mat=mat, mats=sc.materials(), bvh=bvh, ds=ds, cam=sc.cam)
    ds.close()


def _gpu(torch, ds, cam, mode, seeds, points):
    n = int(cam[10]) * int(cam[11])
    rays = torch.empty((8, n), dtype=torch.float32, device="cuda")
    hiprt.camera_rays_device(cam, rays)
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(ds, rays, hits, tracer="while-while")
    fb = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    pdf = torch.empty(n, dtype=torch.float32, device="cuda")
    lc = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    shadow = torch.empty_like(rays)
    sh = hiprt.new_hits(n)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    if mode == "light":
        assert hiprt.set_lights(ds, points) > 0
    for s in seeds:
        hiprt.sample_direct(ds, mode, rays, hits, int(s), shadow, pdf, fb, light_col=lc)
        if mode == "light":
            hiprt.trace_any(ds, shadow, occ, tracer="persistent-while-while")
            hiprt.integrate_light_sample(ds, rays, hits, shadow, occ, pdf, lc, fb)
        else:
            hiprt.trace_closest(ds, shadow, sh, tracer="persistent-while-while")
            hiprt.integrate_dir_sample(ds, rays, hits, shadow, sh, pdf, fb)
    torch.cuda.synchronize()
    return fb.cpu().numpy().reshape(int(cam[11]), int(cam[10]), 4)


@pytest.mark.parametrize("mode,point", [("uniform", False), ("cosine", False), ("light", False), ("light", True)])
def test_direct_modes_match_oracle(torch_cuda, hall, mode, point):
    cam = hall["cam"].copy()
    cam[10], cam[11] = 320, 180
    seeds = orc.frame_seeds(3, 0, 3)
    points = np.array([[*hall["cam"][12:15], 40.0, 35.0, 30.0]], np.float32) if point else None
    print(f"case: direct lighting, hall config {hall['cfg']}, {mode} sampling, point light {point}, 320x180, "
          f"{len(seeds)} samples", flush=True)
    got = _gpu(torch_cuda, hall["ds"], cam, mode, seeds, points)
    b = hall["bvh"]
    want = orc.direct(hall["pos"], hall["idx"], hall["mat"], hall["mats"], b.nodes, b.root, b.index, cam, mode, seeds,
                      points)
    differ = (got.view(np.uint32) != want.view(np.uint32)).any(-1)
    scale = np.maximum(np.abs(want[..., :3]).max(-1), 1e-30)
    far = (np.abs(got[..., :3] - want[..., :3]).max(-1) > PIXEL_REL * scale)
    sums = got[..., :3].sum(dtype=np.float64), want[..., :3].sum(dtype=np.float64)
    print({"pixels_differ": int(differ.sum()), "pixels_beyond_rel": int(far.sum()), "sum_gpu": sums[0],
           "sum_oracle": sums[1]})
    assert np.array_equal(got[..., 3], want[..., 3])
    assert sums[1] > 0
    if mode == "light":
        assert not differ.any()
    else:
        assert far.sum() <= MAX_DIFF * far.size * len(seeds)
        assert abs(sums[0] - sums[1]) <= SUM_REL * sums[1]


def test_point_light_pair_equals_light_mode(torch_cuda, hip_ctx, hall):
    """hiprt_shadow_rays + hiprt_integrate_light (one point light) equal light
    mode over a light list of that point light alone, bit for bit.  Both run
    on a copy of the hall without emission, so the point light is the whole
    list."""
    torch = torch_cuda
    cam = hall["cam"].copy()
    cam[10], cam[11] = 256, 144
    n = 256 * 144
    light, col = hall["cam"][12:15].copy(), np.array([40.0, 35.0, 30.0], np.float32)
    print(f"case: point-light pair vs light mode, hall config {hall['cfg']} without emission, 256x144", flush=True)
    sc = hiprt.HostScene.from_arrays(hall["pos"], hall["idx"])
    ds = hiprt.DeviceScene(hip_ctx, sc, hall["bvh"])
    assert hiprt.set_lights(ds, np.array([[*light, *col]], np.float32)) == 1
    rays = torch.empty((8, n), dtype=torch.float32, device="cuda")
    hiprt.camera_rays_device(cam, rays)
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(ds, rays, hits, tracer="while-while")
    shadow = torch.empty_like(rays)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    fb1 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    hiprt.shadow_rays(ds, rays, hits, shadow, light)
    hiprt.trace_any(ds, shadow, occ, tracer="while-while")
    hiprt.integrate_light(ds, rays, hits, shadow, occ, light, col, fb1)
    fb2 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    pdf = torch.empty(n, dtype=torch.float32, device="cuda")
    lc = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    hiprt.sample_direct(ds, "light", rays, hits, 1, shadow, pdf, fb2, light_col=lc)
    hiprt.trace_any(ds, shadow, occ, tracer="while-while")
    hiprt.integrate_light_sample(ds, rays, hits, shadow, occ, pdf, lc, fb2)
    torch.cuda.synchronize()
    a, b = fb1.cpu().numpy(), fb2.cpu().numpy()
    ds.close()
    assert a[:, :3].sum() > 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("mode", ["uniform", "cosine", "light"])
def test_direct_modes_with_shading_normals_match_oracle(torch_cuda, hip_ctx, hall, mode):
    """Random unit vertex normals: the sample steps flip by the interpolated
    shading normal, the integration weighs by it and a triangle light's cosine
    takes its normal interpolated at the sampled point, as in the oracle."""
    sc = hiprt.HostScene.generate(hall["cfg"])
    rng = np.random.default_rng(12)
    nrm = rng.normal(size=(len(hall["pos"]), 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    sc.set_normals(nrm)
    ds = hiprt.DeviceScene(hip_ctx, sc, hall["bvh"])
    cam = hall["cam"].copy()
    cam[10], cam[11] = 256, 144
    seeds = orc.frame_seeds(4, 0, 2)
    print(f"case: direct lighting with random vertex normals, hall config {hall['cfg']}, {mode}, 256x144", flush=True)
    got = _gpu(torch_cuda, ds, cam, mode, seeds, None)
    ds.close()
    b = hall["bvh"]
    want = orc.direct(hall["pos"], hall["idx"], hall["mat"], hall["mats"], b.nodes, b.root, b.index, cam, mode, seeds,
                      None, normals=nrm)
    plain = orc.direct(hall["pos"], hall["idx"], hall["mat"], hall["mats"], b.nodes, b.root, b.index, cam, mode, seeds,
                       None)
    differ = (got.view(np.uint32) != want.view(np.uint32)).any(-1)
    scale = np.maximum(np.abs(want[..., :3]).max(-1), 1e-30)
    far = (np.abs(got[..., :3] - want[..., :3]).max(-1) > PIXEL_REL * scale)
    print({"pixels_differ": int(differ.sum()), "pixels_beyond_rel": int(far.sum())})
    assert np.array_equal(got[..., 3], want[..., 3])
    assert want[..., :3].sum() > 0
    if mode == "light":
        assert not differ.any()
    else:
        assert far.sum() <= MAX_DIFF * far.size * len(seeds)
    # the normals change most lit pixels
    lit = (want[..., :3].sum(-1) > 0) | (plain[..., :3].sum(-1) > 0)
    assert (plain.view(np.uint32) != want.view(np.uint32)).any(-1)[lit].mean() > 0.3
