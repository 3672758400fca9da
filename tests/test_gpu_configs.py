"""Workload configurations at their stated sizes against the oracle.

* the interior hall (~74k triangles): all 2,073,600 primary rays of the
  1920x1080 camera, every tracer, both tree layouts: the plain and
  persistent-while-while kernels bit-identical to the oracle, the others
  agreeing on >= 99.999% of rays with every mismatch a tie or grazing hit
  (FP64-classified, none unexplained), and t, beta, gamma of agreeing hits no
  further from the FP64 intersection than 1e-6 relative or twice the CPU
  tracer's own error (tests/parity.py:accuracy_ok);
* the 64-scene batch: its smallest and largest scenes (by triangles) and the
  scenes with the most and the fewest rays, at their full ray batches, traced
  together through one descriptor-table launch with the bench's tracer
  options, against the oracle per scene.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
import workloads
from parity import FLT_MAX, accuracy_ok, bit_exact, classify

pytestmark = pytest.mark.gpu


def _np_hits(h):
    return (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(), h["id"].cpu().numpy().view(np.uint32))


@pytest.fixture(scope="module")
def hall2(hip_ctx):
    sc = hiprt.HostScene.generate(2)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    rays = hiprt.camera_rays(sc.cam)
    want = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    ds = {hiprt.BVH2: hiprt.DeviceScene(hip_ctx, sc, bvh), hiprt.BVH4: hiprt.DeviceScene(hip_ctx, sc, bvh, hiprt.BVH4)}
    yield dict(pos=pos, idx=idx, bvh=bvh, rays=rays, want=want, ds=ds)
    for d in ds.values():
        d.close()


C2_CASES = [("while-while", hiprt.BVH2, 0), ("persistent-while-while", hiprt.BVH2, 0),
            ("persistent-while-while", hiprt.BVH2, 4), ("persistent-speculative-while-while", hiprt.BVH2, 0),
            ("dynamic-while-while", hiprt.BVH2, 0), ("dynamic-while-while", hiprt.BVH2, 4),
            ("persistent-while-while", hiprt.BVH4, 3), ("dynamic-while-while", hiprt.BVH4, 0)]


@pytest.mark.parametrize("tracer,layout,stack", C2_CASES)
def test_hall_primary_rays_full_frame(torch_cuda, hall2, tracer, layout, stack):
    h2 = hall2
    n = h2["rays"].shape[1]
    print(f"case: interior hall 1920x1080 primary closest hit, {tracer}, layout {layout}, stack {stack}", flush=True)
    d_rays = torch_cuda.from_numpy(h2["rays"]).cuda()
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(h2["ds"][layout], d_rays, hits, tracer=tracer, stack_config=stack)
    torch_cuda.cuda.synchronize()
    got = _np_hits(hits)
    want = h2["want"]
    assert (want[0] != FLT_MAX).mean() > 0.95
    if layout == hiprt.BVH2 and tracer in ("while-while", "persistent-while-while"):
        assert bit_exact(got, want)
    rep = classify(orc, h2["pos"], h2["idx"], h2["rays"], got, want)
    print(rep)
    assert rep["agree_frac"] >= 0.99999 and rep["unexplained"] == 0
    acc = accuracy_ok(orc, h2["pos"], h2["idx"], h2["rays"], got, want)
    print(acc)
    assert acc["checked"] > 10000 and acc["fail"] == 0


def test_multi_scene_extremes_one_launch(torch_cuda, hip_ctx):
    torch = torch_cuda
    dims = workloads.scene_sizes(64, 4)
    a = np.array(dims)
    picks = sorted({int(a[:, 0].argmin()), int(a[:, 0].argmax()), int(a[:, 1].argmax()), int(a[:, 1].argmin())})
    print(f"case: 64-scene workload, scenes {picks} = {[dims[k] for k in picks]} (tris, rays), one launch", flush=True)
    entries, refs, hosts = [], [], []
    for k in picks:
        sc, bvh, rays = workloads.make_scene(k, dims[k][0], dims[k][1], 4)
        assert rays.shape[1] == dims[k][1]
        ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
        d_rays = torch.from_numpy(rays).cuda()
        h = hiprt.new_hits(rays.shape[1])
        for x in h.values():
            x.fill_(-3)
        entries.append((ds, d_rays, h, None))
        pos, idx, _ = sc.arrays()
        hosts.append((pos, idx, rays))
        refs.append(orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays))
    hiprt.trace_batch(entries, tracer="dynamic-while-while", stack_config=4)
    torch.cuda.synchronize()
    for (ds, _, h, _), want, (pos, idx, rays) in zip(entries, refs, hosts):
        got = _np_hits(h)
        rep = classify(orc, pos, idx, rays, got, want)
        print(rep)
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
        assert rep["agree_frac"] >= 0.99999 and rep["unexplained"] == 0
    for e in entries:
        e[0].close()
