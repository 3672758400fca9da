"""Device refit (new vertex positions, same triangles and topology).

Pinned by: records bit-identical to the host flattener's for the same
vertices (both formats; the Woop rows come from the same FP64 expressions);
node boxes equal to the exact bounding boxes of each subtree's triangles,
computed here from the vertices; refit boxes enclosing the builder's (whose
leaf boxes bound early-split-clipping fragments); and, after the vertices
move, closest and any hits against the oracle's brute-force tracer over the
moved triangles (no BVH involved).
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from bvhcheck import children as _children, subtree_boxes as _subtree_boxes
from parity import FLT_MAX, classify, uniform_sphere_rays

pytestmark = pytest.mark.gpu

TRACERS = list(hiprt.TRACERS)
INVALID = 0x7FFFFFFF


def _moved(pos, seed):
    """Rotate 10 degrees about the y axis through the scene centre, shift,
    and jitter every vertex."""
    rng = np.random.default_rng(seed)
    a = np.deg2rad(10.0)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    c = pos.mean(0)
    out = (pos - c) @ R.T + c + np.array([0.03, -0.02, 0.01]) + rng.normal(0, 0.002, pos.shape)
    return out.astype(np.float32)


@pytest.fixture(scope="module")
def c1r(hip_ctx):
    sc = hiprt.HostScene.generate(1)
    pos, idx, _ = sc.arrays()
    return dict(sc=sc, pos=pos, idx=idx, bvh=hiprt.Bvh(sc))


@pytest.mark.parametrize("fmt", [hiprt.BVH2, hiprt.BVH4])
@pytest.mark.parametrize("tri", [hiprt.TRI_EXACT, hiprt.TRI_WOOP])
def test_refit_boxes_and_records(torch_cuda, hip_ctx, c1r, fmt, tri):
    print(f"case: C1 refit, layout {fmt}, records {tri}", flush=True)
    torch = torch_cuda
    pos, idx, bvh = c1r["pos"], c1r["idx"], c1r["bvh"]
    ds = hiprt.DeviceScene(hip_ctx, c1r["sc"], bvh, fmt, tri)
    n0, r0, s0 = ds.download()
    refs0, lo0, hi0 = _children(n0, fmt)
    # refit to the committed positions
    ds.refit(torch.from_numpy(pos).cuda())
    n1, r1, s1 = ds.download()
    refs1, lo1, hi1 = _children(n1, fmt)
    assert np.array_equal(r1.view(np.uint32), r0.view(np.uint32))
    assert np.array_equal(refs1, refs0) and np.array_equal(s1, s0)
    want_lo, want_hi = _subtree_boxes(refs1, s1, pos, idx)
    assert np.array_equal(lo1, want_lo) and np.array_equal(hi1, want_hi)
    valid = refs0 != INVALID
    slack = 1e-6 * np.maximum(1.0, np.abs(lo0[valid]))
    assert (lo1[valid] <= lo0[valid] + slack).all() and (hi1[valid] >= hi0[valid] - slack).all()
    # move, refit: records equal a host commit of the moved triangles under the same tree
    moved = _moved(pos, 3)
    ds.refit(torch.from_numpy(moved).cuda())
    n2, r2, _ = ds.download()
    host = hiprt.DeviceScene(hip_ctx, hiprt.HostScene.from_arrays(moved, idx), bvh, fmt, tri)
    _, r_host, _ = host.download()
    assert np.array_equal(r2.view(np.uint32), r_host.view(np.uint32))
    refs2, lo2, hi2 = _children(n2, fmt)
    want_lo, want_hi = _subtree_boxes(refs2, s1, moved, idx)
    assert np.array_equal(lo2, want_lo) and np.array_equal(hi2, want_hi)
    # and back: the same bits as the first refit
    ds.refit(torch.from_numpy(pos).cuda())
    n3, r3, _ = ds.download()
    assert np.array_equal(n3, n1) and np.array_equal(r3.view(np.uint32), r1.view(np.uint32))
    host.close()
    ds.close()


@pytest.mark.parametrize("fmt", [hiprt.BVH2, hiprt.BVH4])
def test_traces_after_refit_match_brute_force(torch_cuda, hip_ctx, c1r, fmt):
    print(f"case: C1 moved + refit, layout {fmt}, every tracer vs brute force", flush=True)
    torch = torch_cuda
    pos, idx, bvh = c1r["pos"], c1r["idx"], c1r["bvh"]
    moved = _moved(pos, 4)
    ds = hiprt.DeviceScene(hip_ctx, c1r["sc"], bvh, fmt)
    ds.refit(torch.from_numpy(moved).cuda())
    cam = hiprt.camera_rays(c1r["sc"].cam)
    rays = np.concatenate([cam[:, ::4], uniform_sphere_rays(60_000, 8)], 1)
    want = orc.seq_closest(moved, idx, rays)
    want_any = orc.seq_any(moved, idx, rays)
    d_rays = torch.from_numpy(rays).cuda()
    for tracer in TRACERS[(1 if fmt == hiprt.BVH4 else 0):]:
        h = hiprt.new_hits(rays.shape[1])
        hiprt.trace_closest(ds, d_rays, h, tracer=tracer)
        occ = torch.empty(rays.shape[1], dtype=torch.uint8, device="cuda")
        hiprt.trace_any(ds, d_rays, occ, tracer=tracer)
        torch.cuda.synchronize()
        got = (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(),
               h["id"].cpu().numpy().view(np.uint32))
        rep = classify(orc, moved, idx, rays, got, want)
        assert rep["unexplained"] == 0, (tracer, rep)
        assert rep["agree_frac"] >= 0.99999, (tracer, rep)
        assert (got[0] != FLT_MAX).sum() > 10_000
        assert (occ.cpu().numpy() != want_any).mean() <= 1e-5, tracer
    ds.close()


def test_refit_needs_a_committed_host_scene(torch_cuda, hip_ctx, c1r):
    torch = torch_cuda
    ds = hiprt.DeviceScene(hip_ctx, c1r["sc"], c1r["bvh"])
    nodes, recs, ids = ds.download()
    h = ctypes_upload(hip_ctx, nodes, recs, ids, ds)
    with pytest.raises(hiprt.HiprtError, match="committed from a host scene"):
        h.refit(torch.from_numpy(c1r["pos"]).cuda())
    h.close()
    ds.close()


def ctypes_upload(ctx, nodes, recs, ids, like):
    import ctypes
    out = ctypes.c_void_p()
    hiprt.check(hiprt.lib().hiprt_scene_upload(ctx.h, nodes.ctypes.data, like.n_nodes, like.bvh_format,
                                               recs.ctypes.data, ids.ctypes.data, like.n_slots, like.tri_format,
                                               like.root_ref, ctypes.byref(out)))
    return hiprt.DeviceScene(ctx, None, None, handle=out)
