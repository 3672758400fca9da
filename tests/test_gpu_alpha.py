"""Alpha-aware tracing on the GPU (tracer id | HIPRT_TRACER_ALPHA, the
reference's "<tracer>-alpha" variants) against the oracle's alpha-aware
traversal of the same tree.

Scene: 6000 random triangles in two materials; material 0 carries a 16 x 16
checkerboard texture whose alpha alternates 0.2 / 0.9, material 1 a texture
that is opaque except for one transparent column; texture coordinates are
random in [-2, 3] per vertex (so lookups wrap).  Pinned: the plain and
persistent-while-while kernels bit-identical to the oracle (binary tree); the
postponing kernels and the 4-wide tree identical in t (the closest accepted
hit is unique; only equal-t ties may name another triangle); any-hit bits
identical for every tracer; the non-alpha tracers ignore the textures; the
descriptor-table batch and the path tracer honour the flag.
"""
import numpy as np
import pytest

import hiprt
import oracle_py as orc
from parity import FLT_MAX, bit_exact, uniform_sphere_rays

pytestmark = pytest.mark.gpu

TRACERS = list(hiprt.TRACERS)


def _textures():
    y, x = np.mgrid[0:16, 0:16]
    a = np.ones((16, 16, 4), np.float32)
    a[..., 3] = np.where((x + y) % 2 == 0, 0.2, 0.9)
    b = np.ones((8, 32, 4), np.float32)
    b[:, 5, 3] = 0.0
    return {0: a, 1: b}


@pytest.fixture(scope="module")
def alpha_scene(torch_cuda, hip_ctx):
    base = hiprt.HostScene.random(6000, 0.03, 21)
    pos, idx, _ = base.arrays()
    rng = np.random.default_rng(21)
    mat = (rng.random(len(idx)) < 0.5).astype(np.uint32)
    tc = rng.uniform(-2.0, 3.0, (len(pos), 2)).astype(np.float32)
    sc = hiprt.HostScene.from_arrays(pos, idx, mat)
    sc.set_texcoords(tc)
    tex = _textures()
    for m, img in tex.items():
        sc.set_texture(m, img)
    bvh = hiprt.Bvh(sc)
    ds2 = hiprt.DeviceScene(hip_ctx, sc, bvh)
    ds4 = hiprt.DeviceScene(hip_ctx, sc, bvh, hiprt.BVH4)
    a = orc.Alpha(tc, mat, tex, 2)
    yield dict(sc=sc, pos=pos, idx=idx, bvh=bvh, ds2=ds2, ds4=ds4, alpha=a)
    ds2.close()
    ds4.close()


def _closest(torch, ds, rays, tracer, **kw):
    h = hiprt.new_hits(rays.shape[1])
    hiprt.trace_closest(ds, torch.from_numpy(rays).cuda(), h, tracer=tracer, **kw)
    torch.cuda.synchronize()
    return (h["t"].cpu().numpy(), h["u"].cpu().numpy(), h["v"].cpu().numpy(),
            h["id"].cpu().numpy().view(np.uint32))


def _any(torch, ds, rays, tracer):
    occ = torch.full((rays.shape[1],), 7, dtype=torch.uint8, device="cuda")
    hiprt.trace_any(ds, torch.from_numpy(rays).cuda(), occ, tracer=tracer)
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def test_tracer_names_round_trip():
    for t in TRACERS:
        i = hiprt.tracer_id(t + "-alpha")
        assert i == hiprt.tracer_id(t) | hiprt.TRACER_ALPHA
        assert hiprt.lib().hiprt_tracer_name(i).decode() == t + "-alpha"


@pytest.mark.parametrize("layout", ["bvh2", "bvh4"])
def test_alpha_hits_match_oracle(torch_cuda, alpha_scene, layout):
    torch = torch_cuda
    s = alpha_scene
    rays = uniform_sphere_rays(100_000, 22)
    want = orc.trace_closest_alpha(s["pos"], s["idx"], s["bvh"].nodes, s["bvh"].root, s["bvh"].index, rays,
                                   s["alpha"])
    want_any = orc.trace_any_alpha(s["pos"], s["idx"], s["bvh"].nodes, s["bvh"].root, s["bvh"].index, rays,
                                   s["alpha"])
    plain = orc.trace_closest(s["pos"], s["idx"], s["bvh"].nodes, s["bvh"].root, s["bvh"].index, rays)
    # the textures matter: alpha-aware hits differ from plain ones on many rays
    assert (want[3] != plain[3]).mean() > 0.05
    ds = s["ds2"] if layout == "bvh2" else s["ds4"]
    for tracer in TRACERS:
        if layout == "bvh4" and tracer == "while-while":
            continue
        for stack in (0, 3, 4):
            print(f"case: alpha scene, {layout}, {tracer}-alpha, stack {stack}", flush=True)
            got = _closest(torch, ds, rays, tracer + "-alpha", stack_config=stack)
            if layout == "bvh2" and tracer in ("while-while", "persistent-while-while"):
                assert bit_exact(got, want), (tracer, stack)
            else:
                assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32)), (tracer, stack)
                assert (got[3] != want[3]).mean() <= 1e-4, (tracer, stack)
            if stack == 0:
                assert np.array_equal(_any(torch, ds, rays, tracer + "-alpha"), want_any), tracer


def test_plain_tracers_ignore_textures(torch_cuda, alpha_scene):
    torch = torch_cuda
    s = alpha_scene
    rays = uniform_sphere_rays(30_000, 23)
    want = orc.trace_closest(s["pos"], s["idx"], s["bvh"].nodes, s["bvh"].root, s["bvh"].index, rays)
    for tracer in ("while-while", "persistent-while-while"):
        print(f"case: alpha scene traced by {tracer}", flush=True)
        assert bit_exact(_closest(torch, s["ds2"], rays, tracer), want)


def test_batch_honours_the_alpha_flag(torch_cuda, hip_ctx, alpha_scene):
    torch = torch_cuda
    s = alpha_scene
    other = hiprt.HostScene.generate(1)
    ds_o = hiprt.DeviceScene(hip_ctx, other, hiprt.Bvh(other))   # no textures: alpha changes nothing there
    rays = uniform_sphere_rays(40_000, 24)
    rays_o = uniform_sphere_rays(20_000, 25)
    for tracer in ("dynamic-while-while-alpha", "persistent-while-while"):
        print(f"case: alpha scene + C1 batch, {tracer}", flush=True)
        h, h_o = hiprt.new_hits(rays.shape[1]), hiprt.new_hits(rays_o.shape[1])
        hiprt.trace_batch([(s["ds2"], torch.from_numpy(rays).cuda(), h, None),
                           (ds_o, torch.from_numpy(rays_o).cuda(), h_o, None)], tracer=tracer)
        torch.cuda.synchronize()
        for hh, ds, r in ((h, s["ds2"], rays), (h_o, ds_o, rays_o)):
            want = _closest(torch, ds, r, tracer)
            assert np.array_equal(hh["t"].cpu().numpy().view(np.uint32), want[0].view(np.uint32))
    ds_o.close()


def test_path_tracer_alpha(torch_cuda, alpha_scene):
    s = alpha_scene
    cam = np.array([0.5, 0.5, -1.0, 0, 0, 1, 0, 1, 0, 60, 200, 150], np.float32)
    counts = {}
    for name, batched, tracer in (("plain", False, "while-while-alpha"), ("batched", True, "persistent-while-while-alpha"),
                                  ("batched, no alpha", True, "persistent-while-while")):
        print(f"case: alpha scene path tracer, {name}", flush=True)
        pt = hiprt.PathTracer(s["ds2"], cam, batched=batched, seed=3, tracer=tracer)
        pt.render(0)
        counts[name] = pt.ray_counts().tolist()
        pt.close()
    assert counts["plain"] == counts["batched"], counts
    assert counts["batched"] != counts["batched, no alpha"], counts


def test_obj_texture_coordinates_drive_alpha(torch_cuda, hip_ctx, tmp_path):
    """A textured OBJ (per-corner "vt") traced alpha-aware equals the oracle
    fed with the coordinates the reader assigned."""
    torch = torch_cuda
    base = hiprt.HostScene.random(3000, 0.03, 31)
    pos, idx, _ = base.arrays()
    rng = np.random.default_rng(31)
    uv = rng.uniform(-1.0, 2.0, (3 * len(idx), 2)).astype(np.float32)
    lines = [f"v {x!r} {y!r} {z!r}" for x, y, z in pos.tolist()]
    lines += [f"vt {u!r} {v!r}" for u, v in uv.tolist()]
    lines += [f"f {a + 1}/{3 * t + 1} {b + 1}/{3 * t + 2} {c + 1}/{3 * t + 3}" for t, (a, b, c) in enumerate(idx.tolist())]
    (tmp_path / "tex.obj").write_text("\n".join(lines) + "\n")
    sc = hiprt.HostScene.load_obj(tmp_path / "tex.obj")
    p2, i2, m2 = sc.arrays()
    tc = sc.texcoords()
    assert len(p2) == len(pos) + 3 * len(idx)
    assert np.array_equal(tc[i2.reshape(-1)], uv)   # every corner carries its own coordinate
    tex = _textures()
    sc.set_texture(0, tex[0])
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(hip_ctx, sc, bvh)
    rays = uniform_sphere_rays(50_000, 32)
    a = orc.Alpha(tc, m2, {0: tex[0]}, 1)
    want = orc.trace_closest_alpha(p2, i2, bvh.nodes, bvh.root, bvh.index, rays, a)
    print("case: textured OBJ, persistent-while-while-alpha", flush=True)
    got = _closest(torch, ds, rays, "persistent-while-while-alpha")
    assert bit_exact(got, want)
    plain = orc.trace_closest(p2, i2, bvh.nodes, bvh.root, bvh.index, rays)
    assert (want[3] != plain[3]).mean() > 0.05
    ds.close()


def test_woop_records_alpha(torch_cuda, hip_ctx, alpha_scene):
    """Woop records keep triangle ids in a side array; the opacity test reads
    them there.  Woop barycentrics differ from the exact test's in the last
    bits, so lookups may land on a neighbouring texel for a few rays."""
    torch = torch_cuda
    s = alpha_scene
    ds = hiprt.DeviceScene(hip_ctx, s["sc"], s["bvh"], hiprt.BVH2, hiprt.TRI_WOOP)
    rays = uniform_sphere_rays(50_000, 33)
    want = orc.trace_closest_alpha(s["pos"], s["idx"], s["bvh"].nodes, s["bvh"].root, s["bvh"].index, rays,
                                   s["alpha"])
    for tracer in ("while-while-alpha", "dynamic-while-while-alpha"):
        print(f"case: alpha scene, woop records, {tracer}", flush=True)
        got = _closest(torch, ds, rays, tracer)
        assert (got[3] == want[3]).mean() >= 0.999, tracer
    ds.close()
