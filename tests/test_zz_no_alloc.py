"""The allocation-free closest hit (hiprt_trace_closest_noalloc) against
hiprt_trace_closest: same scene, same rays, same tracer options, bit for bit,
on two calls in a row, the second with fewer rays.  The traces run the
deterministic tracers (plain and persistent-while-while), whose hits do not
depend on the schedule.  The hall's binary tree needs 22 stack entries, so
stack configuration 4 (20 entries) spills two entries per lane into the
scratch's spill column; the random scene's tree fits without spilling.  A
grid whose spill column would not fit the scratch is refused on the host."""
import numpy as np
import pytest

import hiprt
from parity import uniform_sphere_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(hip_ctx):
    out = {}
    for cfg in (1, 3):
        sc = hiprt.HostScene.generate(cfg)
        bvh = hiprt.Bvh(sc)
        out[cfg] = dict(sc=sc, bvh=bvh, ds=hiprt.DeviceScene(hip_ctx, sc, bvh))
    yield out
    for v in out.values():
        v["ds"].close()


def _rays(torch, cfg, scenes, first):
    sc = scenes[cfg]["sc"]
    if first:
        return torch.from_numpy(hiprt.camera_rays(sc.cam)).cuda()
    lo, hi = (0.0, 1.0) if cfg == 1 else (0.5, 14.5)
    return torch.from_numpy(uniform_sphere_rays(300_000, 17, lo=lo, hi=hi)).cuda()


def _np(h):
    return [h[k].cpu().numpy().view(np.uint32) for k in ("t", "u", "v", "id")]


CASES = [(3, "persistent-while-while", 4), (3, "persistent-while-while", 0), (1, "persistent-while-while", 0),
         (1, "while-while", 0), (3, "persistent-while-while", 3)]


@pytest.mark.parametrize("cfg,tracer,stack", CASES)
def test_noalloc_equals_trace_closest(torch_cuda, scenes, cfg, tracer, stack):
    torch = torch_cuda
    ds = scenes[cfg]["ds"]
    need = ds.stack_need
    if cfg == 3 and stack == 4:
        assert need > 20   # the case spills into the scratch's column
    for first in (True, False):
        rays = _rays(torch, cfg, scenes, first)
        n = rays.shape[1]
        print(f"case: allocation-free trace, config {cfg}, {tracer}, stack {stack}, {n} rays, stack need {need}",
              flush=True)
        want = hiprt.new_hits(n)
        hiprt.trace_closest(ds, rays, want, tracer=tracer, stack_config=stack)
        got = hiprt.new_hits(n)
        hiprt.trace_closest_noalloc(ds, rays, got, tracer=tracer, stack_config=stack)
        torch.cuda.synchronize()
        a, b = _np(got), _np(want)
        assert (b[0] != np.float32(3.4028234663852886e38).view(np.uint32)).mean() > 0.3
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_noalloc_refuses_a_spill_column_larger_than_its_scratch(torch_cuda, scenes):
    torch = torch_cuda
    ds = scenes[3]["ds"]
    rays = _rays(torch, 3, scenes, False)
    hits = hiprt.new_hits(rays.shape[1])
    print("case: allocation-free trace, hall, stack 4, 100000 blocks (refused on the host)", flush=True)
    # 100000 blocks x 256 lanes x 2 spill entries x 4 B = 204.8 MB > 16 MiB
    with pytest.raises(hiprt.HiprtError, match="spill column"):
        hiprt.trace_closest_noalloc(ds, rays, hits, stack_config=4, blocks=100000)
    # the next call runs normally
    want = hiprt.new_hits(rays.shape[1])
    hiprt.trace_closest(ds, rays, want, tracer="persistent-while-while", stack_config=4)
    hiprt.trace_closest_noalloc(ds, rays, hits, stack_config=4)
    torch.cuda.synchronize()
    for x, y in zip(_np(hits), _np(want)):
        assert np.array_equal(x, y)
