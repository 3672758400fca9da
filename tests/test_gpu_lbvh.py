"""Device BVH construction (linear BVH, hiprt_scene_build).

Pinned by: the slot order equals a numpy recomputation of the 63-bit Morton
codes (same float32 expressions) sorted stably; every triangle in exactly one
slot; triangle records bit-identical to the host record construction; node
boxes equal to the exact bounds of each subtree's triangles; traces agreeing
with the oracle's brute-force tracer (no BVH involved) for every tracer; the
other entry points (path tracer, refit) working on the result.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from bvhcheck import children, subtree_boxes
from parity import FLT_MAX, classify, uniform_sphere_rays

pytestmark = pytest.mark.gpu

TRACERS = list(hiprt.TRACERS)


def _spread21(x):
    x = x.astype(np.uint64) & np.uint64(0x1FFFFF)
    out = np.zeros_like(x)
    for b in range(21):
        out |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def morton_codes(pos, idx):
    """The build's Morton codes, recomputed with the same float32 steps."""
    tri = pos[idx]
    lo, hi = tri.min(1), tri.max(1)
    c = np.float32(0.5) * (lo + hi)
    blo, bhi = c.min(0), c.max(0)
    ext = bhi - blo
    with np.errstate(divide="ignore"):
        scale = np.where(ext > 0, np.float32(2 ** 21 - 1) / ext, np.float32(0)).astype(np.float32)
    x = ((c - blo) * scale).astype(np.float32)
    q = np.clip(x, 0, np.float32(2 ** 21 - 1)).astype(np.uint32)
    return (_spread21(q[:, 0]) << np.uint64(2)) | (_spread21(q[:, 1]) << np.uint64(1)) | _spread21(q[:, 2])


def _build(torch, ctx, pos, idx, **kw):
    return hiprt.build_device_scene(ctx, torch.from_numpy(np.ascontiguousarray(pos, np.float32)).cuda(),
                                    torch.from_numpy(np.ascontiguousarray(idx).astype(np.int32)).cuda(), **kw)


def _closest(torch, ds, rays, tracer):
    h = hiprt.new_hits(rays.shape[1])
    hiprt.trace_closest(ds, torch.from_numpy(rays).cuda(), h, tracer=tracer)
    torch.cuda.synchronize()
    return (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(),
            h["id"].cpu().numpy().view(np.uint32))


def _any(torch, ds, rays, tracer):
    occ = torch.full((rays.shape[1],), 7, dtype=torch.uint8, device="cuda")
    hiprt.trace_any(ds, torch.from_numpy(rays).cuda(), occ, tracer=tracer)
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def _scenes():
    sc1 = hiprt.HostScene.generate(1)
    pos1, idx1, _ = sc1.arrays()
    rnd = hiprt.HostScene.random(20_000, 0.015, 77)
    pos2, idx2, _ = rnd.arrays()
    # 600 triangles whose boxes share one centre exactly (dyadic offsets
    # c + a, c - a, c + s*a with |s| <= 1): zero extent of the centres, every
    # Morton code 0, the tree decided by sorted position alone
    rng = np.random.default_rng(3)
    a = rng.integers(-64, 65, (600, 3)) / 256.0
    a[a == 0] = 1.0 / 256.0
    s = rng.choice([-0.5, 0.25, 0.5], (600, 3))
    c = 0.5
    pos3 = np.stack([c + a, c - a, c + s * a], 1).reshape(-1, 3).astype(np.float32)
    idx3 = np.arange(pos3.shape[0], dtype=np.uint32).reshape(-1, 3)
    return [("C1", pos1, idx1), ("random 20k", pos2, idx2), ("shared centroid", pos3, idx3)]


@pytest.mark.parametrize("max_leaf", [1, 4])
def test_structure(torch_cuda, hip_ctx, max_leaf):
    torch = torch_cuda
    for name, pos, idx in _scenes():
        print(f"case: device build, {name}, max_leaf {max_leaf}", flush=True)
        ds, ms = _build(torch, hip_ctx, pos, idx, max_leaf=max_leaf)
        nodes, recs, slots = ds.download()
        n = len(idx)
        assert np.array_equal(np.sort(slots), np.arange(n, dtype=np.uint32))
        codes = morton_codes(pos, idx)[slots]
        assert (codes[1:] >= codes[:-1]).all()
        if name == "shared centroid":
            assert (codes == 0).all()
        tie = codes[1:] == codes[:-1]
        assert (slots[1:][tie] > slots[:-1][tie]).all()   # the sort is stable
        # records: (A, id), (A - B), (A - C) in float32
        A, B, C = (pos[idx[slots, k]] for k in range(3))
        want = np.zeros((n, 12), np.float32)
        want[:, 0:3] = A
        want[:, 3] = slots.view(np.float32)
        want[:, 4:7] = A - B
        want[:, 8:11] = A - C
        assert np.array_equal(recs.reshape(n, 12).view(np.uint32), want.view(np.uint32))
        refs, lo, hi = children(nodes, hiprt.BVH2)
        wlo, whi = subtree_boxes(refs, slots, pos, idx)
        assert np.array_equal(lo, wlo) and np.array_equal(hi, whi)
        # leaves hold at most max_leaf triangles, and every slot is in one leaf
        leaf = refs < 0
        cnt = ((~refs[leaf]).astype(np.int64) & 31) + 1
        assert cnt.max() <= max_leaf and cnt.sum() == n
        print(name, {"nodes": int(ds.n_nodes), "depth": int(ds.depth), "build_ms": round(ms, 3)})
        ds.close()


def test_traces_match_brute_force(torch_cuda, hip_ctx):
    torch = torch_cuda
    for name, pos, idx in _scenes()[:2]:
        print(f"case: device build, {name}, every tracer vs brute force", flush=True)
        ds, _ = _build(torch, hip_ctx, pos, idx)
        rays = uniform_sphere_rays(60_000, 5)
        want = orc.seq_closest(pos, idx, rays)
        want_any = orc.seq_any(pos, idx, rays)
        for tracer in TRACERS:
            got = _closest(torch, ds, rays, tracer)
            rep = classify(orc, pos, idx, rays, got, want)
            assert rep["unexplained"] == 0, (name, tracer, rep)
            assert rep["agree_frac"] >= 0.99999, (name, tracer, rep)
            assert (_any(torch, ds, rays, tracer) != want_any).mean() <= 1e-5
        ds.close()


@pytest.mark.parametrize("n_tris", [0, 1, 3])
def test_tiny_and_empty_scenes(torch_cuda, hip_ctx, n_tris):
    print(f"case: device build of {n_tris} triangles", flush=True)
    torch = torch_cuda
    rng = np.random.default_rng(n_tris)
    pos = (rng.uniform(0.3, 0.7, (n_tris, 1, 3)) + rng.normal(0, 0.8, (n_tris, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    ds, _ = _build(torch, hip_ctx, pos, idx)
    assert ds.n_nodes == 0
    rays = uniform_sphere_rays(5000, 7)
    if n_tris:
        want = orc.seq_closest(pos, idx, rays)
    else:
        want = (np.full(5000, FLT_MAX, np.float32), np.zeros(5000, np.float32), np.zeros(5000, np.float32),
                np.zeros(5000, np.uint32))
    for tracer in TRACERS:
        got = _closest(torch, ds, rays, tracer)
        rep = classify(orc, pos, idx, rays, got, want) if n_tris else {"unexplained": 0, "agree_frac": 1.0}
        assert rep["unexplained"] == 0 and rep["agree_frac"] >= 0.9999, (tracer, rep)
        if not n_tris:
            assert (got[0] == FLT_MAX).all()
    ds.close()


def test_rejects_bad_input(torch_cuda, hip_ctx):
    torch = torch_cuda
    pos = np.random.default_rng(1).uniform(0, 1, (30, 3)).astype(np.float32)
    idx = np.arange(30, dtype=np.uint32).reshape(-1, 3)
    bad = idx.copy()
    bad[4, 1] = 30
    with pytest.raises(hiprt.HiprtError, match="past n_vert"):
        _build(torch, hip_ctx, pos, bad)
    with pytest.raises(hiprt.HiprtError, match="max_leaf"):
        _build(torch, hip_ctx, pos, idx, max_leaf=33)
    mats = torch.full((10,), 2, dtype=torch.int32, device="cuda")
    with pytest.raises(hiprt.HiprtError, match="material id"):
        _build(torch, hip_ctx, pos, idx, materials=mats, material_table=np.zeros((2, 8), np.float32))


def test_hall_path_tracer_and_refit(torch_cuda, hip_ctx):
    print("case: device build of the hall, path tracer plain vs batched, refit", flush=True)
    torch = torch_cuda
    sc = hiprt.HostScene.generate(3)
    pos, idx, mat = sc.arrays()
    table = sc.materials()
    assert table.shape[0] > int(mat.max())
    ds_host = hiprt.DeviceScene(hip_ctx, sc, hiprt.Bvh(sc))
    ds, ms = _build(torch, hip_ctx, pos, idx, materials=torch.from_numpy(mat.astype(np.int32)).cuda(),
                    material_table=table)
    print({"hall_build_ms": round(ms, 3), "nodes": ds.n_nodes, "depth": ds.depth})
    assert ds.depth <= 32   # fits the default persistent stack
    cam = sc.cam.copy()
    cam[10], cam[11] = 320, 180
    imgs = []
    for dsc, batched in ((ds, False), (ds, True), (ds_host, False)):
        pt = hiprt.PathTracer(dsc, cam, batched=batched, seed=4,
                              tracer="persistent-while-while" if batched else None)
        pt.render(0)
        imgs.append(pt.image())
        pt.close()
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    # against the host-built tree: the same paths except where a grazing hit or
    # a tie picks another triangle
    same = (imgs[0] == imgs[2]).all(axis=-1)
    assert same.mean() >= 0.999, same.mean()
    # refit to the same vertices leaves the device-built tree unchanged: its
    # leaf boxes already bound whole triangles
    n0, r0, s0 = ds.download()
    ds.refit(torch.from_numpy(pos).cuda())
    n1, r1, s1 = ds.download()
    assert np.array_equal(n0, n1) and np.array_equal(r0.view(np.uint32), r1.view(np.uint32))
    ds.close()
    ds_host.close()
