"""Traversals deeper than a stack configuration's VGPR and LDS entries.

120 nested triangles, the k-th on the plane x + y + z = 2^-k (jittered),
give a binned-SAH tree 38 levels deep: the stack configurations hold fewer
entries in registers and LDS than its traversals need (the 4-wide tree's
need exceeds all four, the binary tree's at least configurations 1-3), so
the oldest entries go to the global spill column.  Pinned by: the
oracle's traversal of the same binary tree (bit-identical for the tracer that
keeps the reference's order), the oracle's brute-force tracer (no BVH) for
every persistent tracer, stack configuration and tree width, the batched
entry point and the path tracer on the same scene.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, bit_exact, classify

pytestmark = pytest.mark.gpu

PERSISTENT = [t for t in hiprt.TRACERS if t != "while-while"]
N_TRIS = 120


def nested_scene():
    s = 0.5 ** np.arange(N_TRIS)
    tri = np.zeros((N_TRIS, 3, 3))
    tri[:, 0, 0] = s
    tri[:, 1, 1] = s
    tri[:, 2, 2] = s
    tri += 0.01 * s[:, None, None] * np.random.default_rng(2).uniform(-1, 1, (N_TRIS, 3, 3))
    pos = tri.reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * N_TRIS, dtype=np.uint32).reshape(-1, 3)
    return pos, idx


def nested_rays(n, seed):
    """Origins at every nesting depth (log-uniform scale 1 .. 2^-40), uniform
    directions."""
    rng = np.random.default_rng(seed)
    scale = 2.0 ** -rng.uniform(0, 40, n)
    o = rng.uniform(0, 1, (n, 3)) * scale[:, None]
    z = 1.0 - 2.0 * rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.empty((8, n), np.float32)
    rays[0:3] = o.T
    rays[3] = 0.0
    rays[4:7] = d.T
    rays[7] = FLT_MAX
    return rays


@pytest.fixture(scope="module")
def deep(torch_cuda, hip_ctx):
    pos, idx = nested_scene()
    sc = hiprt.HostScene.from_arrays(pos, idx)
    bvh = hiprt.Bvh(sc, max_tris=1)
    ds2 = hiprt.DeviceScene(hip_ctx, sc, bvh)
    ds4 = hiprt.DeviceScene(hip_ctx, sc, bvh, hiprt.BVH4)
    print({"depth": bvh.depth, "need_bvh2": ds2.stack_need, "need_bvh4": ds4.stack_need}, flush=True)
    yield pos, idx, sc, bvh, ds2, ds4
    ds2.close()
    ds4.close()


def _closest(torch, ds, rays, tracer, **kw):
    h = hiprt.new_hits(rays.shape[1])
    hiprt.trace_closest(ds, torch.from_numpy(rays).cuda(), h, tracer=tracer, **kw)
    torch.cuda.synchronize()
    return (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(),
            h["id"].cpu().numpy().view(np.uint32))


def _any(torch, ds, rays, tracer, **kw):
    occ = torch.full((rays.shape[1],), 7, dtype=torch.uint8, device="cuda")
    hiprt.trace_any(ds, torch.from_numpy(rays).cuda(), occ, tracer=tracer, **kw)
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def test_tree_outgrows_every_stack_configuration(deep):
    *_, ds2, ds4 = deep
    # configurations hold 32 (4 VGPR + 28 LDS), 24, 22 and 24 entries
    assert ds2.stack_need > 24 and ds4.stack_need > 32


@pytest.mark.parametrize("layout", ["bvh2", "bvh4"])
@pytest.mark.parametrize("tracer", PERSISTENT)
def test_spilled_traversal_matches_oracle(torch_cuda, deep, tracer, layout):
    torch = torch_cuda
    pos, idx, sc, bvh, ds2, ds4 = deep
    ds = ds2 if layout == "bvh2" else ds4
    rays = nested_rays(20_000, 11)
    want = orc.seq_closest(pos, idx, rays)
    want_any = orc.seq_any(pos, idx, rays)
    same_tree = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    assert (want[0] < FLT_MAX).mean() > 0.3
    for stack in range(5):
        print(f"case: nested scene, {layout}, {tracer}, stack {stack}", flush=True)
        got = _closest(torch, ds, rays, tracer, stack_config=stack)
        if layout == "bvh2" and tracer == "persistent-while-while":
            assert bit_exact(got, same_tree), stack
        rep = classify(orc, pos, idx, rays, got, want)
        assert rep["unexplained"] == 0 and rep["agree_frac"] >= 0.99999, (stack, rep)
        occ = _any(torch, ds, rays, tracer, stack_config=stack)
        assert (occ != want_any).mean() <= 1e-4, stack


def _check_lost(torch, ds, rays, stack, tracer="persistent-while-while"):
    """Capped and uncapped runs of the same rays: lost rays are marked, the
    others equal (t, beta, gamma, id for the reference-order tracer, t for a
    speculative one)."""
    full = _closest(torch, ds, rays, tracer, stack_config=stack)
    capped = _closest(torch, ds, rays, tracer, stack_config=stack, spill_limit=1)
    lost = np.isnan(capped[0])
    print({"lost": int(lost.sum()), "rays": rays.shape[1]})
    assert lost.any() and (~lost).mean() > 0.3
    assert np.all(capped[3][lost] == 0xFFFFFFFF) and not np.isnan(full[0]).any()
    # the speculative tracers may return either triangle of an equal-t tie
    exact = tracer == "persistent-while-while"
    for a, b in zip(capped if exact else capped[:1], full if exact else full[:1]):
        assert np.array_equal(a[~lost].view(np.uint32), b[~lost].view(np.uint32))
    occ_full = _any(torch, ds, rays, tracer, stack_config=stack)
    occ_cap = _any(torch, ds, rays, tracer, stack_config=stack, spill_limit=1)
    lost_any = occ_cap == 0xFF
    assert set(np.unique(occ_full).tolist()) <= {0, 1}
    assert np.array_equal(occ_cap[~lost_any], occ_full[~lost_any])


def test_capped_spill_column_marks_lost_rays_4wide(torch_cuda, deep):
    """hiprt_tracer_opts.spill_limit = 1: a traversal that outgrows the stack
    configuration's entries plus one spill entry drops entries (never writing
    out of bounds); its result is marked lost -- t NaN, id 0xFFFFFFFF,
    occlusion byte 0xFF -- and every other ray equals the uncapped run bit for
    bit.  The nested scene's 4-wide tree with stack configuration 3, its
    random rays plus rays on the plane x + y + z = 0 through the nesting's
    corner (every box holds that corner, no triangle reaches the plane, so
    these traversals hold the most entries)."""
    torch = torch_cuda
    pos, idx, sc, bvh, ds2, ds4 = deep
    rng = np.random.default_rng(14)
    graze = np.zeros((8, 2000), np.float32)
    a = 2.0 ** -rng.uniform(5, 40, 2000)
    graze[0], graze[1], graze[2] = -a, a, 0.0
    d = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    graze[4], graze[5], graze[6] = d[0], d[1], d[2]
    graze[7] = FLT_MAX
    rays = np.ascontiguousarray(np.concatenate([nested_rays(20_000, 13), graze], 1))
    print("case: nested scene, bvh4, persistent-while-while, stack 3, spill_limit 1", flush=True)
    _check_lost(torch, ds4, rays, 3)


def test_spilled_batch_and_work_chunks(torch_cuda, hip_ctx, deep):
    torch = torch_cuda
    pos, idx, sc, bvh, ds2, ds4 = deep
    rays = nested_rays(30_000, 12)
    other = hiprt.HostScene.generate(1)
    ds_o = hiprt.DeviceScene(hip_ctx, other, hiprt.Bvh(other))
    pos_o, idx_o, _ = other.arrays()
    from parity import uniform_sphere_rays
    rays_o = uniform_sphere_rays(25_000, 13)
    for stack in (1, 3, 4):
        for chunks in (1, 8):
            print(f"case: nested + C1 batch, stack {stack}, chunks {chunks}", flush=True)
            d_r, d_o = torch.from_numpy(rays).cuda(), torch.from_numpy(rays_o).cuda()
            h, h_o = hiprt.new_hits(rays.shape[1]), hiprt.new_hits(rays_o.shape[1])
            hiprt.trace_batch([(ds2, d_r, h, None), (ds_o, d_o, h_o, None)], stack_config=stack,
                              work_chunks=chunks)
            torch.cuda.synchronize()
            for hh, ds, r in ((h, ds2, rays), (h_o, ds_o, rays_o)):
                want = _closest(torch, ds, r, "dynamic-while-while", stack_config=stack, work_chunks=chunks)
                got = (hh["t"].cpu().numpy(), hh["u"].cpu().numpy(), hh["v"].cpu().numpy(),
                       hh["id"].cpu().numpy().view(np.uint32))
                # one launch or two: only ties between equal t may differ
                diff = got[3] != want[3]
                assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
                assert diff.mean() <= 1e-4
    ds_o.close()


def test_spilled_path_tracer(torch_cuda, deep):
    torch = torch_cuda
    pos, idx, sc, bvh, ds2, ds4 = deep
    # camera near the apex of the nesting, looking at it
    cam = np.array([0.02, 0.02, 0.02, -1, -1, -1, 0, 1, 0, 60, 160, 120], np.float32)
    imgs = {}
    for name, ds, batched, stack, srt in (("plain", ds2, False, 0, False), ("batched bvh2", ds2, True, 3, False),
                                          ("batched bvh2 stack 4", ds2, True, 4, True),
                                          ("batched bvh4", ds4, True, 3, True)):
        print(f"case: nested scene path tracer, {name}", flush=True)
        pt = hiprt.PathTracer(ds, cam, batched=batched, seed=5, stack_config=stack, sort_rays=srt,
                              tracer="persistent-while-while" if batched else None)
        pt.render(0)
        imgs[name] = pt.image()
        pt.close()
    assert np.array_equal(imgs["plain"].view(np.uint32), imgs["batched bvh2"].view(np.uint32))
    assert np.array_equal(imgs["plain"].view(np.uint32), imgs["batched bvh2 stack 4"].view(np.uint32))
    same = (imgs["plain"] == imgs["batched bvh4"]).all(axis=-1)
    assert same.mean() >= 0.999, same.mean()
