"""GPU wavefront steps and the path tracer.

* device camera rays are bit-identical to the host (and oracle) camera rays;
* device cosine bounce rays match the oracle's restatement of
  sample_cos_weighted_dir driven by the same PCG32 numbers: hit points to the
  bit, directions to 1e-5 (cosf/sinf differ between libm and the device by an
  ulp, which the hemisphere's square root amplifies near grazing directions);
* the diffuse rays of the hall trace bit-identically on GPU and CPU;
* the batched (compacted, persistent, graph-replayed) path tracer produces
  the same per-level ray counts and the same image as the plain per-bounce
  one, with bounce queues sorted by ray key or in shading order (a ray
  scattered to the wrong slot would continue another pixel's path), and its
  level counts follow the Russian-roulette law.  These bit-for-bit checks run
  the persistent-while-while tracer, whose hits never depend on the schedule;
  the speculative and dynamic tracers may pick the other triangle of an
  equal-t tie and are checked against the oracle with its tolerance
  (tests/test_gpu_pt_oracle.py).
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, bit_exact
from pcg import Pcg32, stream_id

pytestmark = pytest.mark.gpu


def _hits_np(h):
    return (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(), h["id"].cpu().numpy().view(np.uint32))


@pytest.fixture(scope="module")
def hall(hip_ctx):
    sc = hiprt.HostScene.generate(3)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
    return dict(sc=sc, pos=pos, idx=idx, bvh=bvh, ds=ds, cam=sc.cam)


def test_device_camera_rays_equal_host(torch_cuda, hall):
    print("case: hall camera 1920x1080, device vs host camera rays", flush=True)
    want = hiprt.camera_rays(hall["cam"])
    out = torch_cuda.empty(want.shape, dtype=torch_cuda.float32, device="cuda")
    hiprt.camera_rays_device(hall["cam"], out)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))


def test_incoherent_rays_match_oracle_and_trace(torch_cuda, hall):
    """hiprt_incoherent_rays (the camera step's "incoherence" rays,
    rt/cuda/kernels.cu:75-156) on the hall's spheres: the same PCG32 streams
    as the oracle, so points agree to the ulps of the device's sinf / cosf /
    acosf / cbrtf (5e-6 of the scene's extent), zero radii keep the centres
    exactly, t spans [0, FLT_MAX]; the rays trace to the oracle's hits
    except where those ulps move a ray across an edge."""
    print("case: hall, 1M incoherent rays r1 = 0.5 r2 = 0.3, device vs oracle", flush=True)
    torch = torch_cuda
    s1, s2, r_max = hiprt.incoherence_spheres(hall["pos"])
    n = 1 << 20
    out = torch.empty((8, n), dtype=torch.float32, device="cuda")
    hiprt.incoherent_rays_device(s1, s2, 0.5, 0.3, r_max, 1234, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = orc.incoherent_rays(s1, s2, 0.5, 0.3, r_max, 1234, n)
    scale = float(r_max) * 4
    assert np.abs(got[0:3] - want[0:3]).max() <= 5e-6 * scale
    assert np.abs(got[4:7] - want[4:7]).max() <= 1e-5 * scale
    assert (got[3] == 0).all() and (got[7] == FLT_MAX).all()
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(hall["ds"], out, hits, tracer="while-while")
    torch.cuda.synchronize()
    t, _, _, ids = _hits_np(hits)
    bvh = hall["bvh"]
    tw, _, _, iw = orc.trace_closest(hall["pos"], hall["idx"], bvh.nodes, bvh.root, bvh.index, got[:, :65536])
    assert np.array_equal(ids[:65536], iw) and np.array_equal(t[:65536].view(np.uint32), tw.view(np.uint32))
    assert (t < FLT_MAX).mean() > 0.9
    zero = torch.empty((8, 256), dtype=torch.float32, device="cuda")
    hiprt.incoherent_rays_device(s1, s2, 0.0, 0.0, r_max, 1, zero)
    z = zero.cpu().numpy()
    assert (z[0:3].T == s1).all() and (z[4:7].T == s2 - s1).all()
    with pytest.raises(hiprt.HiprtError, match="non-negative"):
        hiprt.incoherent_rays_device(s1, s2, -0.5, 0.0, r_max, 1, zero)


def test_bounce_rays_match_oracle_and_trace_exactly(torch_cuda, hall):
    print("case: hall 480x270 primary hits -> cosine bounce -> diffuse trace", flush=True)
    torch = torch_cuda
    cam = hall["cam"].copy()
    cam[10], cam[11] = 480, 270
    rays = hiprt.camera_rays(cam)
    n = rays.shape[1]
    d_rays = torch.from_numpy(rays).cuda()
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(hall["ds"], d_rays, hits, tracer="while-while")
    seed, level = 77, 0
    out = torch.empty_like(d_rays)
    hiprt.bounce_cosine(hall["ds"], d_rays, hits, out, seed, level)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    h = _hits_np(hits)
    rng = Pcg32(stream_id(seed, np.arange(n, dtype=np.uint64), level))
    xi0, xi1 = rng.next_float(), rng.next_float()
    want = orc.cosine_rays(hall["pos"], hall["idx"], rays, *h, xi0, xi1)
    assert np.array_equal(got[[0, 1, 2, 3, 7]].view(np.uint32), want[[0, 1, 2, 3, 7]].view(np.uint32))
    # cosf/sinf differ by an ulp or two between libm and the device, and the
    # hemisphere's z = sqrt(1 - x^2 - y^2) amplifies that near the rim
    assert np.max(np.abs(got[4:7] - want[4:7])) < 1e-5
    assert np.median(np.abs(got[4:7] - want[4:7])) < 1e-7
    # the device's own diffuse rays through GPU and CPU tracers
    hits2 = hiprt.new_hits(n)
    hiprt.trace_closest(hall["ds"], out, hits2, tracer="while-while")
    torch.cuda.synchronize()
    ref = orc.trace_closest(hall["pos"], hall["idx"], hall["bvh"].nodes, hall["bvh"].root, hall["bvh"].index, got)
    assert bit_exact(_hits_np(hits2), ref)


def test_bounce_rays_with_shading_normals_match_oracle(torch_cuda, hip_ctx, hall):
    """With per-vertex shading normals (random unit vectors, so that about
    half the hits flip differently from the geometric rule) the bounce step
    flips by the interpolated shading normal, as the oracle does."""
    print("case: hall 480x270 primary hits -> cosine bounce with random vertex normals", flush=True)
    torch = torch_cuda
    sc = hiprt.HostScene.generate(3)
    pos, idx, _ = sc.arrays()
    rng = np.random.default_rng(5)
    nrm = rng.normal(size=(len(pos), 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    sc.set_normals(nrm)
    ds = hiprt.DeviceScene(hip_ctx, sc, hall["bvh"])
    cam = hall["cam"].copy()
    cam[10], cam[11] = 480, 270
    rays = hiprt.camera_rays(cam)
    n = rays.shape[1]
    d_rays = torch.from_numpy(rays).cuda()
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(ds, d_rays, hits, tracer="while-while")
    out = torch.empty_like(d_rays)
    hiprt.bounce_cosine(ds, d_rays, hits, out, 91, 2)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ds.close()
    h = _hits_np(hits)
    g = Pcg32(stream_id(91, np.arange(n, dtype=np.uint64), 2))
    xi0, xi1 = g.next_float(), g.next_float()
    want = orc.cosine_rays(pos, idx, rays, *h, xi0, xi1, normals=nrm)
    plain = orc.cosine_rays(pos, idx, rays, *h, xi0, xi1)
    assert np.array_equal(got[[0, 1, 2, 3, 7]].view(np.uint32), want[[0, 1, 2, 3, 7]].view(np.uint32))
    assert np.max(np.abs(got[4:7] - want[4:7])) < 1e-5
    # the normals change the side of a large share of the bounces
    flipped = (np.sign(want[4:7]) != np.sign(plain[4:7])).any(0) & (want[7] > 0)
    assert flipped.mean() > 0.2


@pytest.mark.parametrize("res,stack,chunks,sort", [((320, 180), 0, 1, True), ((320, 180), 1, 8, True),
                                                   ((320, 180), 0, 8, False), ((333, 97), 3, 8, True),
                                                   ((333, 97), 3, 1, False), ((333, 97), 4, 8, True)])
def test_batched_equals_plain(torch_cuda, hall, res, stack, chunks, sort):
    print(f"case: hall path tracer {res}, plain vs batched (stack {stack}, chunks {chunks}, sorted queues {sort}), "
          "2 frames", flush=True)
    cam = hall["cam"].copy()
    cam[10], cam[11] = res
    plain = hiprt.PathTracer(hall["ds"], cam, batched=False, seed=5)
    batched = hiprt.PathTracer(hall["ds"], cam, batched=True, seed=5, stack_config=stack, work_chunks=chunks,
                               sort_rays=sort, tracer="persistent-while-while")
    counts = []
    for pt in (plain, batched):
        pt.render(0)
        c0 = pt.ray_counts()
        pt.render(1)
        counts.append((c0, pt.ray_counts()))
    img_p, img_b = plain.image(), batched.image()
    plain.close()
    batched.close()
    n = res[0] * res[1]
    for (a0, a1), name in zip(counts, ("plain", "batched")):
        print(name, a0.tolist(), a1.tolist())
    assert counts[0][0].tolist() == counts[1][0].tolist()
    assert counts[0][1].tolist() == counts[1][1].tolist()
    c = counts[1][0]
    assert c[0] == n and c[1] <= n and c[2] <= c[1]
    # Russian roulette with p = 0.8 at levels 3 and 4: survivors / previous
    # level is 0.8 times the (closed hall: ~1) hit fraction, binomial 6 sigma
    assert c[1] >= 0.99 * n and c[2] >= 0.98 * c[1]
    for lvl in (3, 4):
        p_hat = c[lvl] / c[lvl - 1]
        sd = np.sqrt(0.8 * 0.2 / c[lvl - 1])
        assert 0.8 * 0.98 - 6 * sd <= p_hat <= 0.8 + 6 * sd
    assert np.isfinite(img_p).all() and img_p.sum() > 0
    # identical paths, identical accumulation order per pixel
    assert np.array_equal(img_p.view(np.uint32), img_b.view(np.uint32))


@pytest.mark.parametrize("res,depth", [((13, 7), 4), ((64, 40), 0), ((64, 40), 1), ((47, 45), 2)])
def test_small_frames_and_depths(torch_cuda, hall, res, depth):
    """Queues shorter than one sort tile, no bounce at all, and a single bounce:
    the sorted batched schedule still equals the plain one."""
    print(f"case: hall path tracer {res}, max_depth {depth}, plain vs sorted batched", flush=True)
    cam = hall["cam"].copy()
    cam[10], cam[11] = res
    imgs, counts = [], []
    for batched in (False, True):
        pt = hiprt.PathTracer(hall["ds"], cam, max_depth=depth, batched=batched, seed=11,
                              tracer="persistent-while-while" if batched else None)
        pt.render(0)
        pt.render(1)
        imgs.append(pt.image())
        counts.append(pt.ray_counts().tolist())
        pt.close()
    assert counts[0] == counts[1]
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))


@pytest.mark.parametrize("samples,res", [(3, (320, 180)), (2, (61, 33))])
def test_samples_in_flight(torch_cuda, hall, samples, res):
    """samples > 1: one render call traces the samples of `samples`
    consecutive frames together.  Per level the same rays as rendering those
    frames one by one (ray totals equal), and the same image."""
    print(f"case: hall path tracer {res}, {samples} samples in flight vs one at a time", flush=True)
    cam = hall["cam"].copy()
    cam[10], cam[11] = res
    one = hiprt.PathTracer(hall["ds"], cam, batched=True, seed=13, tracer="persistent-while-while")
    for f in range(2 * samples):
        one.render(f)
    many = hiprt.PathTracer(hall["ds"], cam, batched=True, seed=13, samples=samples, tracer="persistent-while-while")
    many.render(0)
    many.render(1)
    assert one.ray_totals().tolist() == many.ray_totals().tolist()
    a, b = one.image(), many.image()
    one.close()
    many.close()
    assert a.sum() > 0
    # a path adds its radiance once (at the emitter that ends it), so summing
    # per sample first changes no bits
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_graph_replays_are_stable(torch_cuda, hall):
    print("case: hall path tracer 256x144, batched graph replayed 4x vs eager", flush=True)
    cam = hall["cam"].copy()
    cam[10], cam[11] = 256, 144
    a = hiprt.PathTracer(hall["ds"], cam, batched=True, seed=9, tracer="persistent-while-while")
    b = hiprt.PathTracer(hall["ds"], cam, batched=False, seed=9)
    imgs = []
    for pt in (a, b):
        for f in range(4):
            pt.render(f)
            torch_cuda.cuda.synchronize()
            _ = torch_cuda.ones(1000, device="cuda").sum().item()   # eager work between replays
        imgs.append(pt.image())
    a.close()
    b.close()
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
