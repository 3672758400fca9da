"""Closest-hit adjoint on the GPU: the hall (262,144 triangles), 1M seeded
diffuse rays (cosine bounces of the 1920x1080 primary hits, seed 5), loss
L = sum_i a_i t_i + b_i beta_i + c_i gamma_i with a, b, c ~ N(0, 1).

The fp32 backward (d L / d vertices, origins, directions; hit set frozen) is
checked on 10,000 sampled rays against the FP64 analytic gradient and FP64
central differences (h = 1e-4) of the oracle, and the torch.autograd wrapper
against the kernel's own per-ray outputs.  Hit ids are never differentiated.

Tolerances are relative to the per-ray gradient's largest FP64 component:
fp32 vs FP64 analytic 1e-3 on 99.9% of rays (5e-3 on 99.99%); FP64 analytic
vs FP64 central differences 1e-3 on 99% of rays (central differences lose
accuracy only on near-grazing rays, where t changes fast).
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5(hip_ctx, torch_cuda):
    torch = torch_cuda
    sc = hiprt.HostScene.generate(3)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
    cam = torch.empty((8, 1920 * 1080), dtype=torch.float32, device="cuda")
    hiprt.camera_rays_device(sc.cam, cam)
    hits = hiprt.new_hits(cam.shape[1])
    hiprt.trace_closest(ds, cam, hits, tracer="while-while")
    diff = torch.empty_like(cam)
    hiprt.bounce_cosine(ds, cam, hits, diff, seed=5, level=0)
    rays = diff[:, :1_000_000].contiguous()
    torch.cuda.synchronize()
    return dict(torch=torch, ds=ds, pos=pos, idx=idx, rays=rays)


def _rel(a, b):
    scale = np.maximum(np.abs(b).max(axis=1), 1e-12)
    return np.abs(a - b).max(axis=1) / scale


def test_adjoint_vs_fp64_and_central_differences(c5):
    print("case: C5 adjoint, hall, 1M diffuse rays, 10k sampled", flush=True)
    torch, ds, rays = c5["torch"], c5["ds"], c5["rays"]
    n = rays.shape[1]
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(ds, rays, hits, tracer="dynamic-while-while")
    rng = np.random.default_rng(5)
    g = rng.normal(size=(3, n)).astype(np.float32)
    gt, gu, gv = (torch.from_numpy(x).cuda() for x in g)
    corners = torch.empty((n, 9), dtype=torch.float32, device="cuda")
    go = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    gd = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    nv, _ = hiprt.scene_counts(ds)
    gvert = torch.zeros((nv, 3), dtype=torch.float32, device="cuda")
    hiprt.adjoint_closest(ds, rays, hits, gt, gu, gv, grad_vertices=gvert, grad_origin=go, grad_dir=gd,
                          grad_corners=corners)
    torch.cuda.synchronize()
    t = hits["t"].cpu().numpy()
    tri = hits["id"].cpu().numpy().view(np.uint32)
    valid = np.nonzero(t != np.float32(3.4028234663852886e38))[0]
    assert len(valid) > 0.99 * n
    sample = np.sort(rng.choice(valid, 10_000, replace=False))
    r_np = rays.cpu().numpy()[:, sample]
    gpu = np.concatenate([corners.cpu().numpy()[sample], go.cpu().numpy()[sample], gd.cpu().numpy()[sample]], 1)
    args = (c5["pos"], c5["idx"], tri[sample], r_np, g[0][sample], g[1][sample], g[2][sample])
    exact = orc.adjoint_fp64(*args)
    fd = orc.adjoint_fp64(*args, fd_step=1e-4)
    e_gpu, e_fd = _rel(gpu, exact), _rel(exact, fd)
    print({"gpu_vs_fp64_p50": float(np.median(e_gpu)), "p99.9": float(np.quantile(e_gpu, 0.999)),
           "max": float(e_gpu.max()), "fd_p50": float(np.median(e_fd)), "fd_p99": float(np.quantile(e_fd, 0.99))})
    assert (e_gpu <= 1e-3).mean() >= 0.999
    assert (e_gpu <= 5e-3).mean() >= 0.9999
    assert (e_fd <= 1e-3).mean() >= 0.99
    # misses get zero gradients; translation invariance dA+dB+dC+do = 0
    miss = np.nonzero(t == np.float32(3.4028234663852886e38))[0]
    if len(miss):
        assert np.all(go.cpu().numpy()[miss] == 0) and np.all(gd.cpu().numpy()[miss] == 0)
    s = gpu[:, 0:3] + gpu[:, 3:6] + gpu[:, 6:9] + gpu[:, 9:12]
    assert np.max(np.abs(s) / np.maximum(np.abs(gpu).max(1, keepdims=True), 1e-12)) < 1e-4
    # vertex gradients: the per-ray corner gradients scattered to shared vertices
    cn = corners.cpu().numpy().astype(np.float64)
    want = np.zeros((nv, 3))
    vidx = c5["idx"][tri[valid]]
    for k in range(3):
        np.add.at(want, vidx[:, k], cn[valid, 3 * k:3 * k + 3])
    got = gvert.cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-4 * max(1.0, np.abs(want).max())


def test_autograd_function(c5):
    print("case: torch.autograd wrapper of the closest hit, 65,536 rays", flush=True)
    import hiprt_autograd
    torch, ds = c5["torch"], c5["ds"]
    rays = c5["rays"][:, :65536]
    verts = torch.from_numpy(c5["pos"]).cuda().requires_grad_(True)
    o = rays[0:3].t().contiguous().requires_grad_(True)
    d = rays[4:7].t().contiguous().requires_grad_(True)
    t, u, v, tri = hiprt_autograd.closest_hit(ds, verts, o, d)
    assert not tri.requires_grad
    valid = t < 3.0e38
    rng = np.random.default_rng(7)
    a, b, c = (torch.from_numpy(rng.normal(size=t.shape[0]).astype(np.float32)).cuda() for _ in range(3))
    zero = torch.zeros_like(t)
    loss = (a * torch.where(valid, t, zero) + b * torch.where(valid, u, zero) + c * torch.where(valid, v, zero)).sum()
    loss.backward()
    # the same gradients straight from the kernel
    hits = {"t": t.detach(), "u": u.detach(), "v": v.detach(), "id": tri}
    go = torch.empty((t.shape[0], 3), dtype=torch.float32, device="cuda")
    gd = torch.empty_like(go)
    gv = torch.zeros_like(verts)
    z = torch.zeros_like(a)
    hiprt.adjoint_closest(ds, rays.contiguous(), hits, torch.where(valid, a, z), torch.where(valid, b, z),
                          torch.where(valid, c, z), grad_vertices=gv, grad_origin=go, grad_dir=gd)
    torch.cuda.synchronize()
    assert torch.equal(o.grad, go) and torch.equal(d.grad, gd)
    # vertex gradients are sums of per-ray terms added by float atomics, whose
    # order varies between launches: equal up to the rounding of the sums
    diff = float((verts.grad - gv).abs().max())
    scale = float(gv.abs().max())
    print({"vertex_grad_max_diff": diff, "scale": scale})
    assert scale > 0 and diff <= 1e-4 * max(1.0, scale)
    # hits reported by the wrapper are the plain trace's
    h2 = hiprt.new_hits(t.shape[0])
    hiprt.trace_closest(ds, rays.contiguous(), h2, tracer="persistent-while-while")
    assert torch.equal(h2["id"], tri) and torch.equal(h2["t"], t.detach())


@pytest.mark.parametrize("n", [1, 257, 5000, 70_001])
def test_vertex_gradients_accumulate_by_triangle(c5, n):
    """The vertex gradients (rays grouped by hit triangle, summed per run,
    added to the vertices) equal the per-ray corner gradients scattered to
    the shared vertices in float64, and are added to what the array held.
    Every third ray gets t_max just above t_min, so it misses."""
    print(f"case: adjoint vertex accumulation, {n} rays, every third a miss", flush=True)
    torch, ds = c5["torch"], c5["ds"]
    rays = c5["rays"][:, :n].contiguous()
    rays[7, ::3] = 2e-4
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(ds, rays, hits, tracer="persistent-while-while")
    rng = np.random.default_rng(n)
    gt, gu, gv = (torch.from_numpy(rng.normal(size=n).astype(np.float32)).cuda() for _ in range(3))
    nv, _ = hiprt.scene_counts(ds)
    gvert = torch.ones((nv, 3), dtype=torch.float32, device="cuda")
    corners = torch.empty((n, 9), dtype=torch.float32, device="cuda")
    hiprt.adjoint_closest(ds, rays, hits, gt, gu, gv, grad_vertices=gvert, grad_corners=corners)
    torch.cuda.synchronize()
    t = hits["t"].cpu().numpy()
    tri = hits["id"].cpu().numpy().view(np.uint32)
    valid = np.nonzero(t != np.float32(3.4028234663852886e38))[0]
    assert len(valid) < n or n == 1
    cn = corners.cpu().numpy().astype(np.float64)
    want = np.ones((nv, 3))
    vidx = c5["idx"][tri[valid]]
    for k in range(3):
        np.add.at(want, vidx[:, k], cn[valid, 3 * k:3 * k + 3])
    got = gvert.cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.abs(want).max())
