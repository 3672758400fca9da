"""GPU parity of the gfx950 trace kernels against the CPU oracle.

Every case traces the same rays through the same BVH on the GPU (through the
C ABI) and in the oracle.  Exact triangle records with the plain kernel must
be bit-identical to the oracle; the persistent kernels may differ only where
two triangles have equal t; Woop records are held to the tie/grazing bar of
tests/parity.py and to the CPU tracer's own accuracy against FP64.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, bit_exact, classify, fp64_error, uniform_sphere_rays

pytestmark = pytest.mark.gpu

TRACERS = list(hiprt.TRACERS)


def _run_closest(torch, dscene, rays, tracer, **kw):
    r = torch.from_numpy(rays).cuda()
    n = rays.shape[1]
    hits = hiprt.new_hits(n)
    for x in hits.values():
        x.fill_(-7)
    hiprt.trace_closest(dscene, r, hits, tracer=tracer, **kw)
    torch.cuda.synchronize()
    return (hits["t"].cpu().numpy(), hits["u"].cpu().numpy(), hits["v"].cpu().numpy(),
            hits["id"].cpu().numpy().view(np.uint32))


def _run_any(torch, dscene, rays, tracer):
    r = torch.from_numpy(rays).cuda()
    occ = torch.full((rays.shape[1],), 7, dtype=torch.uint8, device="cuda")
    hiprt.trace_any(dscene, r, occ, tracer=tracer)
    torch.cuda.synchronize()
    return occ.cpu().numpy()


@pytest.fixture(scope="module")
def c1(hip_ctx):
    sc = hiprt.HostScene.generate(1)
    pos, idx, _ = sc.arrays()
    bvh = hiprt.Bvh(sc)
    exact = hiprt.DeviceScene(hip_ctx, sc, bvh, hiprt.BVH2, hiprt.TRI_EXACT)
    woop = hiprt.DeviceScene(hip_ctx, sc, bvh, hiprt.BVH2, hiprt.TRI_WOOP)
    cam_rays = hiprt.camera_rays(sc.cam)
    rnd = uniform_sphere_rays(100_000, seed=11)
    rays = np.concatenate([cam_rays, rnd], axis=1)
    ref = orc.trace_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays)
    return dict(sc=sc, pos=pos, idx=idx, bvh=bvh, exact=exact, woop=woop, rays=rays, ref=ref)


def test_plain_exact_is_bit_identical(torch_cuda, c1):
    print("case: C1 camera+random rays, while-while, exact records", flush=True)
    got = _run_closest(torch_cuda, c1["exact"], c1["rays"], "while-while")
    assert bit_exact(got, c1["ref"])


@pytest.mark.parametrize("tracer", TRACERS[1:])
def test_persistent_exact_matches_up_to_ties(torch_cuda, c1, tracer):
    print(f"case: C1 camera+random rays, {tracer}, exact records", flush=True)
    got = _run_closest(torch_cuda, c1["exact"], c1["rays"], tracer)
    rep = classify(orc, c1["pos"], c1["idx"], c1["rays"], got, c1["ref"])
    assert rep["unexplained"] == 0, rep
    assert rep["agree_frac"] >= 0.99999, rep
    # where the triangle agrees, so do t, u, v, to the bit
    same = (got[3] == c1["ref"][3]) & (got[0] != FLT_MAX)
    for a, b in zip(got[:3], c1["ref"][:3]):
        assert np.array_equal(a[same].view(np.uint32), b[same].view(np.uint32))


@pytest.mark.parametrize("stack", [1, 2, 3, 4])
@pytest.mark.parametrize("tracer", TRACERS[1:])
def test_other_stack_configs_match_up_to_ties(torch_cuda, c1, tracer, stack):
    """Stack configurations 1-4 (register / LDS splits of the per-lane
    stack, and configuration 4's register cap) change only where entries
    live, never the traversal order."""
    print(f"case: C1 camera+random rays, {tracer}, exact records, stack config {stack}", flush=True)
    got0 = _run_closest(torch_cuda, c1["exact"], c1["rays"], tracer)
    got1 = _run_closest(torch_cuda, c1["exact"], c1["rays"], tracer, stack_config=stack)
    rep = classify(orc, c1["pos"], c1["idx"], c1["rays"], got1, c1["ref"])
    assert rep["unexplained"] == 0, rep
    same_t = got0[0].view(np.uint32) == got1[0].view(np.uint32)
    assert same_t.all()


@pytest.mark.parametrize("chunks,blocks", [(8, 0), (8, 3), (1, 0), (1, 3)])
@pytest.mark.parametrize("tracer", TRACERS[1:])
def test_work_chunks_match_up_to_ties(torch_cuda, c1, tracer, chunks, blocks):
    """One work counter, or eight chunks (one per XCD, the default) also with
    fewer blocks than chunks (every block then serves several chunks): every
    ray traced once, same hits up to ties."""
    print(f"case: C1 camera+random rays, {tracer}, {chunks} work chunks, blocks {blocks}", flush=True)
    got = _run_closest(torch_cuda, c1["exact"], c1["rays"], tracer, work_chunks=chunks, blocks=blocks)
    assert (got[0].view(np.uint32) != np.float32(-7).view(np.uint32)).all()   # every hit slot written
    rep = classify(orc, c1["pos"], c1["idx"], c1["rays"], got, c1["ref"])
    assert rep["unexplained"] == 0, rep
    assert rep["agree_frac"] >= 0.99999, rep


@pytest.mark.parametrize("tracer", ["while-while", "dynamic-while-while"])
def test_woop_within_tie_and_grazing_bar(torch_cuda, c1, tracer):
    """Woop records: hit agreement >= 99.999% and every mismatch a tie or a
    grazing hit, and their accuracy against FP64 no worse than the reference's
    own fp32 Shirley test on the same hits (the oracle's hits, which the exact
    records reproduce bit for bit): at least the CPU tracer's fraction of t
    within 1e-6 relative, t's largest error within twice the CPU's, and beta's
    and gamma's 99th-percentile errors within twice the CPU's.  The records
    hold the triangle's corner as a local origin (core/records.h), so no row
    carries an offset term whose rounding grows with |A|; per ray the two
    methods round differently, so the comparison is over the distribution."""
    print(f"case: C1 camera+random rays, {tracer}, woop records", flush=True)
    got = _run_closest(torch_cuda, c1["woop"], c1["rays"], tracer)
    rep = classify(orc, c1["pos"], c1["idx"], c1["rays"], got, c1["ref"])
    print(rep)
    assert rep["unexplained"] == 0, rep
    assert rep["agree_frac"] >= 0.99999, rep
    # the same rays for both: those where the two hit the same triangle
    tw, uw, vw, iw = got
    tr, ur, vr, ir = c1["ref"]
    both = np.nonzero((tw != FLT_MAX) & (tr != FLT_MAX) & (iw == ir))[0]
    both = np.random.default_rng(3).choice(both, min(50000, len(both)), replace=False)
    sub = lambda h: tuple(x[both] for x in h)
    _, et, eu, ev = fp64_error(orc, c1["pos"], c1["idx"], c1["rays"][:, both], sub(got))
    _, ct, cu, cv = fp64_error(orc, c1["pos"], c1["idx"], c1["rays"][:, both], sub(c1["ref"]))
    q = lambda x: float(np.quantile(x, 0.99))
    stats = {"t_frac_1e-6": float((et <= 1e-6).mean()), "cpu_t_frac_1e-6": float((ct <= 1e-6).mean()),
             "t_max": float(et.max()), "cpu_t_max": float(ct.max()), "u_p99": q(eu), "cpu_u_p99": q(cu),
             "v_p99": q(ev), "cpu_v_p99": q(cv), "u_max": float(eu.max()), "v_max": float(ev.max())}
    print(stats)
    assert stats["t_frac_1e-6"] >= stats["cpu_t_frac_1e-6"]
    assert stats["t_max"] <= 2 * stats["cpu_t_max"]
    assert stats["u_p99"] <= 2 * stats["cpu_u_p99"] and stats["v_p99"] <= 2 * stats["cpu_v_p99"]
    assert stats["u_max"] <= 1e-3 and stats["v_max"] <= 1e-3


@pytest.mark.parametrize("tracer", TRACERS)
def test_any_hit_bits_identical(torch_cuda, c1, tracer):
    print(f"case: C1 shadow rays, any hit, {tracer}", flush=True)
    rays = c1["rays"].copy()
    rays[7] = np.where(np.arange(rays.shape[1]) % 3 == 0, np.float32(0.7), rays[7])  # short and long rays
    want = orc.trace_any(c1["pos"], c1["idx"], c1["bvh"].nodes, c1["bvh"].root, c1["bvh"].index, rays)
    got = _run_any(torch_cuda, c1["exact"], rays, tracer)
    assert np.array_equal(got, want)
