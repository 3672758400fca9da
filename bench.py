#!/usr/bin/env python3
"""Benchmark: closest-hit Mrays/s of the 4-bounce diffuse path tracer on the
262,144-triangle hall at 1920x1080 -- batched (persistent kernel, compacted
queues, hipGraph replay) against plain per-bounce launches.

One step = one path-traced frame: camera rays plus up to four cosine-weighted
diffuse bounces (Russian roulette p = 0.8 at bounces 3 and 4), every level's
rays traced by the gfx950 closest-hit kernel.  `value` counts every traced
ray of the timed frames over their wall time on the GPU stream (inputs -- the
flattened BVH and triangle records -- already resident in HBM).

Prints one JSON line.  Also reported: the plain schedule's rate and the ratio,
the trace kernels alone on primary and on diffuse rays, the roofline of the
dominant kernel (algorithmic bytes per launch from the counting kernel
variant, time from HIP events on the launch stream) and the CPU oracle's
throughput on a bounded sample of the same diffuse rays.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s closest-hit, 262k-tri scene, diffuse rays: batched vs plain kernel"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_frames(torch, pt, steps, warmup, stream):
    for f in range(warmup):
        pt.render(f, stream.cuda_stream)
    torch.cuda.synchronize()
    before = pt.ray_totals()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for f in range(steps):
        pt.render(warmup + f, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    after = pt.ray_totals()
    ms = e0.elapsed_time(e1) / steps
    rays = float((after - before).sum()) / steps
    return ms, rays, (after - before) / steps


def time_trace(torch, hiprt, ds, rays, tracer="dynamic-while-while", stream=None, reps=10, **kw):
    n = rays.shape[1]
    hits = hiprt.new_hits(n)
    for _ in range(2):
        hiprt.trace_closest(ds, rays, hits, tracer=tracer, stream=stream.cuda_stream, **kw)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        hiprt.trace_closest(ds, rays, hits, tracer=tracer, stream=stream.cuda_stream, **kw)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, hits


def kernel_name(kw, ds, binary):
    """Template instantiation of the persistent closest-hit kernel that a
    path-tracer configuration runs on scene `ds` (exact records)."""
    tracer = kw.get("tracer", "dynamic-while-while")
    spec = "true" if tracer in ("persistent-speculative-while-while", "dynamic-while-while") else "false"
    refill = "true" if tracer == "dynamic-while-while" else "false"
    reg = {0: 4, 1: 8, 2: 2, 3: 0, 4: 0}[kw.get("stack_config", 0)]
    lds = {0: 28, 1: 16, 2: 20, 3: 24, 4: 20}[kw.get("stack_config", 0)]
    # the LDS ring with a spill column when the tree outgrows the
    # configuration, the plain LDS stack otherwise
    ring = "true" if ds.stack_need > reg + lds else "false"
    return (f"k_trace_persistent<0, false, {spec}, {refill}, {reg}, {lds}, false, {2 if binary else 4}, false, "
            f"{ring}> ({tracer}, stack {kw.get('stack_config', 0)}, {'binary' if binary else '4-wide'} tree)")


def time_calls(torch, fn, stream, reps=5):
    fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_shadow(torch, hiprt, ctx, stream):
    """Hall of 74,068 triangles at 1920x1080: primary closest hits, then one
    shadow ray per hit to the point light (t_max = distance - 1e-4), traced
    with any hit."""
    sc = hiprt.HostScene.generate(2)
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(ctx, sc, bvh)
    cam = sc.cam
    n = int(cam[10]) * int(cam[11])
    prim = torch.empty((8, n), dtype=torch.float32, device="cuda")
    hiprt.camera_rays_device(cam, prim, stream.cuda_stream)
    hits = hiprt.new_hits(n)
    hiprt.trace_closest(ds, prim, hits, tracer="while-while", stream=stream.cuda_stream)
    shadow = torch.empty_like(prim)
    hiprt.shadow_rays(ds, prim, hits, shadow, cam[12:15], stream.cuda_stream)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda")
    out = {"triangles": sc.n_tris, "rays": n}
    for tr, kw in (("while-while", {}), ("dynamic-while-while", {}), ("dynamic-while-while", {"work_chunks": 1}),
                   ("dynamic-while-while", {"stack_config": 4}), ("persistent-speculative-while-while", {}),
                   ("persistent-speculative-while-while", {"stack_config": 4})):
        ms = time_calls(torch, lambda: hiprt.trace_any(ds, shadow, occ, tracer=tr, stream=stream.cuda_stream, **kw),
                        stream)
        tag = "".join(f",{k}={v}" for k, v in kw.items())
        out[f"any/{tr}{tag}"] = round(n / ms / 1e3, 1)
    ms = time_calls(torch, lambda: hiprt.trace_closest(ds, prim, hits, tracer="dynamic-while-while",
                                                      stream=stream.cuda_stream), stream)
    out["primary/dynamic-while-while"] = round(n / ms / 1e3, 1)
    out["occluded_frac"] = round(float(occ.float().mean()), 4)
    ds.close()
    return out


def bench_device_build(torch, hiprt, ctx, stream, sc, d_diff, host_build_s):
    """Linear BVH of the hall built on the device from device-resident
    triangles, and the diffuse trace over it (the tree is not the SAH one)."""
    pos, idx, mat = sc.arrays()
    d_pos = torch.from_numpy(pos).cuda()
    d_idx = torch.from_numpy(idx.astype("int32")).cuda()
    d_mat = torch.from_numpy(mat.astype("int32")).cuda()
    table = sc.materials()
    runs = []
    for _ in range(3):
        t0 = time.time()
        ds, ms = hiprt.build_device_scene(ctx, d_pos, d_idx, materials=d_mat, material_table=table)
        runs.append((ms, time.time() - t0))
        if len(runs) < 3:
            ds.close()
    ms_trace, _ = time_trace(torch, hiprt, ds, d_diff, "dynamic-while-while", stream)
    out = {"triangles": int(len(idx)), "device_ms": round(min(r[0] for r in runs), 3),
           "call_ms": round(1e3 * min(r[1] for r in runs), 1), "host_sah_esc_s": round(host_build_s, 3),
           "inner_nodes": int(ds.n_nodes), "depth": int(ds.depth),
           "diffuse_mrays": round(d_diff.shape[1] / ms_trace / 1e3, 1)}
    ds.close()
    return out


def bench_alpha(torch, hiprt, ctx, stream, d_cam, d_diff):
    """Alpha-aware tracing on the hall: every material gets a 16 x 16 texture
    whose alpha alternates 0.2 / 0.9 (a cut-out checkerboard), texture
    coordinates from the vertex position (x / 2, z / 2).  The same rays are
    traced with the plain and the alpha-aware dynamic tracer."""
    import numpy as np
    sc = hiprt.HostScene.generate(3)
    pos, _, _ = sc.arrays()
    sc.set_texcoords(np.stack([pos[:, 0] * 0.5, pos[:, 2] * 0.5], 1))
    y, x = np.mgrid[0:16, 0:16]
    img = np.ones((16, 16, 4), np.float32)
    img[..., 3] = np.where((x + y) % 2 == 0, 0.2, 0.9)
    for m in range(len(sc.materials())):
        sc.set_texture(m, img)
    ds = hiprt.DeviceScene(ctx, sc, hiprt.Bvh(sc))
    out = {"texture": "16x16 checkerboard alpha 0.2/0.9 on every material"}
    for name, rays in (("primary", d_cam), ("diffuse", d_diff)):
        for tracer in ("dynamic-while-while", "dynamic-while-while-alpha"):
            for kw in ({}, {"stack_config": 4}):
                ms, _ = time_trace(torch, hiprt, ds, rays, tracer, stream, **kw)
                tag = "".join(f",{k}={v}" for k, v in kw.items())
                out[f"{name}/{tracer}{tag}"] = round(rays.shape[1] / ms / 1e3, 1)
    ds.close()
    return out


def bench_multi_scene(torch, hiprt, ctx, stream, n_scenes=64):
    """64 random-triangle scenes of log-uniform size with their own rays: one
    descriptor-table launch against one launch per scene, both with the
    dynamic tracer and stack configuration 4; each scene's layout (binary or
    4-wide) picked by the autotuner on a sample of its rays."""
    import workloads
    t0 = time.time()
    work = workloads.multi_scene(n_scenes, seed=4)
    build_s = time.time() - t0
    entries, scenes, layouts = [], [], []
    total = 0
    n_tris = sum(sc.n_tris for sc, _, _ in work)
    # the same 64 trees built on the device instead (linear BVHs)
    dev_ms = 0.0
    for sc, _, _ in work:
        pos, idx, _ = sc.arrays()
        dsd, ms = hiprt.build_device_scene(ctx, torch.from_numpy(pos).cuda(),
                                           torch.from_numpy(idx.astype("int32")).cuda())
        dev_ms += ms
        dsd.close()
    for sc, bvh, rays in work:
        d_rays = torch.from_numpy(rays).cuda()
        sample = d_rays[:, ::max(1, rays.shape[1] // 262144)].contiguous()
        res, ds = hiprt.autotune(ctx, sc, bvh, sample)
        layouts.append(res["bvh_format"])
        hits = hiprt.new_hits(rays.shape[1])
        entries.append((ds, d_rays, hits, None))
        scenes.append(ds)
        total += rays.shape[1]
    del work
    tune_s = time.time() - t0 - build_s   # includes the device builds above

    def separate():
        for ds, r, h, _ in entries:
            hiprt.trace_closest(ds, r, h, tracer="dynamic-while-while", stack_config=4, stream=stream.cuda_stream)

    def batched():
        hiprt.trace_batch(entries, tracer="dynamic-while-while", stack_config=4, stream=stream.cuda_stream)

    ms_sep = time_calls(torch, separate, stream, reps=3)
    ms_bat = time_calls(torch, batched, stream, reps=3)
    out = {"scenes": n_scenes, "triangles": int(n_tris), "rays": int(total),
           "batched_mrays": round(total / ms_bat / 1e3, 1), "separate_mrays": round(total / ms_sep / 1e3, 1),
           "batched_ms": round(ms_bat, 3), "separate_ms": round(ms_sep, 3),
           "bvh4_scenes": sum(1 for l in layouts if l == "bvh4"),
           "host_build_s": round(build_s, 1), "device_build_ms": round(dev_ms, 1), "autotune_s": round(tune_s, 1)}
    for ds in scenes:
        ds.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the device-build, shadow-ray, 64-scene and alpha legs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (calibrates HBM byte counters)")
    ap.add_argument("--kernel-only", choices=("primary", "diffuse"),
                    help="only trace the hall's primary or first-bounce diffuse rays 3 times with the dynamic "
                         "tracer (for counter collection) and print their fetch counts")
    ap.add_argument("--stack-config", type=int, default=0, help="stack configuration for --kernel-only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (rays are independent, so "
              "the path shards by image tiles across GPUs; see the sharding paragraph of the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "illum-rt_amd"))
    import hiprt

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    stream = torch.cuda.current_stream()
    t0 = time.time()
    ctx = hiprt.Context(0)
    sc = hiprt.HostScene.generate(3)
    pos, idx, _ = sc.arrays()
    t_build = time.time()
    bvh = hiprt.Bvh(sc)
    host_build_s = time.time() - t_build
    ds = hiprt.DeviceScene(ctx, sc, bvh)
    cam = sc.cam
    w, h = int(cam[10]), int(cam[11])
    log(f"scene: {len(idx)} triangles, {ds.n_nodes} inner nodes, {ds.n_slots} leaf slots, depth {ds.depth}; "
        f"setup {time.time() - t0:.1f}s")

    # Camera rays and first-bounce diffuse rays of the frame.
    d_cam = torch.empty((8, w * h), dtype=torch.float32, device="cuda")
    hiprt.camera_rays_device(cam, d_cam, stream.cuda_stream)
    cam_hits = hiprt.new_hits(w * h)
    hiprt.trace_closest(ds, d_cam, cam_hits, tracer="while-while", stream=stream.cuda_stream)
    d_diff = torch.empty_like(d_cam)
    hiprt.bounce_cosine(ds, d_cam, cam_hits, d_diff, seed=12345, level=0, stream=stream.cuda_stream)
    torch.cuda.synchronize()

    if args.kernel_only:
        rays = d_cam if args.kernel_only == "primary" else d_diff
        ms, _ = time_trace(torch, hiprt, ds, rays, "dynamic-while-while", stream, reps=3,
                           stack_config=args.stack_config)
        st = torch.zeros(2, dtype=torch.int64, device="cuda")
        hiprt.trace_closest(ds, rays, hiprt.new_hits(rays.shape[1]), tracer="dynamic-while-while",
                            stream=stream.cuda_stream, fetch_stats=st, stack_config=args.stack_config)
        torch.cuda.synchronize()
        print(json.dumps({"rays": int(rays.shape[1]), "kind": args.kernel_only, "ms": round(ms, 4),
                          "inner_nodes": int(st[0]), "triangle_tests": int(st[1])}), flush=True)
        ds.close()
        ctx.close()
        return

    # Autotune the persistent tracer on the frame's first-bounce diffuse rays
    # (tracer, refill threshold, stack configuration, work chunks, BVH
    # layout); the batched schedule runs with the chosen configuration.
    t_tune = time.time()
    tuned, ds_tuned = hiprt.autotune(ctx, sc, bvh, d_diff)
    tuned["seconds"] = round(time.time() - t_tune, 1)
    log(f"autotuned: {tuned}")
    pt_kw = dict(tracer=tuned["tracer"], stack_config=tuned["stack_config"], work_chunks=tuned["work_chunks"],
                 refill_threshold=tuned["refill_threshold"], blocks=tuned["blocks"])

    # The trace autotuner times one launch over rays already in pixel order
    # with the tree warm in L2; inside a frame every level follows a shading
    # pass and reads a sorted queue.  So the frame's configuration is chosen
    # on warm-up frames (never the timed ones), in rays per second: the best
    # of the autotuned one, its one-counter variant, the 4-wide and the binary
    # tree with stack configuration 3, the defaults with 2, 4 or 8 samples per
    # pixel in flight (consecutive frames' samples traced together, so every
    # launch holds 2-8x the rays) and the binary tree with stack configuration
    # 4 (a 20-entry LDS stack with registers capped for 8 waves per SIMD) and 8,
    # 16, 32 or 64 samples in flight, when it beats
    # the library defaults (binary tree, dynamic tracer, stack 0, eight work
    # chunks, one sample) by more than 2%, the defaults otherwise, which keeps
    # the choice stable between runs.  The autotuner's sample is the first
    # bounce; the deeper bounces weigh the stack configurations and layouts
    # differently, so the fixed candidates are tried on whole frames too.
    ds4 = hiprt.DeviceScene(ctx, sc, bvh, hiprt.BVH4)
    candidates = {"autotuned": (ds_tuned, pt_kw),
                  "autotuned, one work counter": (ds_tuned, dict(pt_kw, work_chunks=1)),
                  "4-wide tree, stack 3": (ds4, dict(tracer="dynamic-while-while", stack_config=3)),
                  "binary tree, stack 3": (ds, dict(tracer="dynamic-while-while", stack_config=3)),
                  "2 samples in flight": (ds, dict(samples=2)),
                  "4 samples in flight": (ds, dict(samples=4)),
                  "8 samples in flight": (ds, dict(samples=8)),
                  "8 samples in flight, stack 4": (ds, dict(tracer="dynamic-while-while", stack_config=4,
                                                            samples=8)),
                  "16 samples in flight, stack 4": (ds, dict(tracer="dynamic-while-while", stack_config=4,
                                                             samples=16)),
                  "32 samples in flight, stack 4": (ds, dict(tracer="dynamic-while-while", stack_config=4,
                                                             samples=32)),
                  "64 samples in flight, stack 4": (ds, dict(tracer="dynamic-while-while", stack_config=4,
                                                             samples=64)),
                  "library defaults": (ds, {})}
    # Rates, not frame times: a render call with k samples in flight traces k
    # samples per pixel.
    trial = {}
    for name, (dsc, kw) in candidates.items():
        pt = hiprt.PathTracer(dsc, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True, **kw)
        ms, rays, _ = time_frames(torch, pt, 4, 1, stream)
        trial[name] = rays / ms / 1e3
        pt.close()
    best_tuned = max((k for k in candidates if k != "library defaults"), key=trial.get)
    choice = best_tuned if trial[best_tuned] > 1.02 * trial["library defaults"] else "library defaults"
    one_sample = max(v for k, v in trial.items() if "samples" not in candidates[k][1])
    log(f"frame configuration trial (Mrays/s): {trial}; timed frames use {choice}")
    ds_frame, frame_kw = candidates[choice]
    samples = frame_kw.get("samples", 1)

    batched = hiprt.PathTracer(ds_frame, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True,
                               **frame_kw)
    plain = hiprt.PathTracer(ds, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=False)
    ms_b, rays_b, per_level_b = time_frames(torch, batched, args.steps, args.warmup, stream)
    log(f"batched: {ms_b:.3f} ms/frame, {rays_b / 1e6:.3f} Mrays/frame, per level {per_level_b.tolist()}")
    ms_p, rays_p, per_level_p = time_frames(torch, plain, args.steps, args.warmup, stream)
    log(f"plain:   {ms_p:.3f} ms/frame, {rays_p / 1e6:.3f} Mrays/frame")
    value = rays_b / ms_b / 1e3
    plain_value = rays_p / ms_p / 1e3
    # Dominant kernel: the persistent closest-hit trace.  Per level: HIP-event
    # time of the launch, and the counting variant's fetches on the same rays.
    node_bytes = 64.0 if (ds_frame.bvh_format == hiprt.BVH2) else 128.0
    ms_lvl, fetch, rays_lvl = batched.profile(frame=10_000, stream=stream.cuda_stream)
    bytes_lvl = 48.0 * rays_lvl + node_bytes * fetch[:, 0] + 48.0 * fetch[:, 1]
    achieved = float(bytes_lvl.sum()) / (float(ms_lvl.sum()) * 1e-3) / 1e9
    launches = int((rays_lvl > 0).sum())
    footprint = ds_frame.n_nodes * node_bytes + ds_frame.n_slots * 48
    compulsory = float(48.0 * rays_lvl.sum() + footprint * launches) / (float(ms_lvl.sum()) * 1e-3) / 1e9
    log(f"trace per level ms {np.round(ms_lvl, 3).tolist()}, nodes/ray "
        f"{np.round(fetch[:, 0] / np.maximum(rays_lvl, 1), 1).tolist()}, tris/ray "
        f"{np.round(fetch[:, 1] / np.maximum(rays_lvl, 1), 1).tolist()}")

    # The trace kernels alone: primary rays and first-bounce diffuse rays.
    kernels = {}
    variants = (("while-while", {}), ("dynamic-while-while", {}),
                ("dynamic-while-while", {"stack_config": 1}), ("dynamic-while-while", {"stack_config": 2}),
                ("dynamic-while-while", {"stack_config": 3}), ("dynamic-while-while", {"stack_config": 4}),
                ("dynamic-while-while", {"work_chunks": 1}),
                ("dynamic-while-while", {"stack_config": 4, "work_chunks": 1}))
    for name, rays in (("primary", d_cam), ("diffuse", d_diff)):
        for tr, kw in variants:
            ms, _ = time_trace(torch, hiprt, ds, rays, tr, stream, **kw)
            tag = "".join(f",{k}={v}" for k, v in kw.items())
            kernels[f"{name}/{tr}{tag}"] = round(rays.shape[1] / ms / 1e3, 1)
        ms, _ = time_trace(torch, hiprt, ds_tuned, rays, stream=stream, **pt_kw)
        kernels[f"{name}/autotuned"] = round(rays.shape[1] / ms / 1e3, 1)
    log(f"trace kernels alone (Mrays/s): {kernels}")

    # Closest-hit adjoint on 1M of the diffuse rays: vertex (atomic), origin
    # and direction gradients of sum(a t + b beta + c gamma), hit set frozen.
    n_adj = 1_000_000
    adj_rays = d_diff[:, :n_adj].contiguous()
    adj_hits = hiprt.new_hits(n_adj)
    hiprt.trace_closest(ds, adj_rays, adj_hits, tracer="dynamic-while-while", stream=stream.cuda_stream)
    gen = torch.Generator(device="cuda").manual_seed(5)
    g_t, g_u, g_v = (torch.randn(n_adj, device="cuda", generator=gen) for _ in range(3))
    g_vert = torch.zeros((hiprt.scene_counts(ds)[0], 3), dtype=torch.float32, device="cuda")
    g_o = torch.empty((n_adj, 3), dtype=torch.float32, device="cuda")
    g_d = torch.empty((n_adj, 3), dtype=torch.float32, device="cuda")
    ms_adj = time_calls(torch, lambda: hiprt.adjoint_closest(ds, adj_rays, adj_hits, g_t, g_u, g_v, g_vert, g_o, g_d,
                                                             stream=stream.cuda_stream), stream, reps=10)
    adjoint = {"rays": n_adj, "ms": round(ms_adj, 4), "mrays": round(n_adj / ms_adj / 1e3, 1)}
    log(f"adjoint: {adjoint}")

    # Device refit of the hall to its own vertices (records and boxes rebuilt;
    # leaf boxes then bound whole triangles instead of clipped fragments), and
    # the diffuse trace over the refitted tree.  Runs last on this scene.
    d_pos = torch.from_numpy(pos).cuda()
    ms_refit = time_calls(torch, lambda: ds.refit(d_pos, stream.cuda_stream), stream, reps=10)
    ms_after, _ = time_trace(torch, hiprt, ds, d_diff, "dynamic-while-while", stream)
    refit = {"ms": round(ms_refit, 4), "levels_launched": int(ds.depth),
             "diffuse_after_refit_mrays": round(d_diff.shape[1] / ms_after / 1e3, 1)}
    log(f"refit: {refit}")

    extra = {}
    if not args.no_extra:
        extra["device_build"] = bench_device_build(torch, hiprt, ctx, stream, sc, d_diff, host_build_s)
        log(f"device build: {extra['device_build']}")
        extra["shadow_any_hit_mrays"] = bench_shadow(torch, hiprt, ctx, stream)
        log(f"shadow rays: {extra['shadow_any_hit_mrays']}")
        extra["multi_scene"] = bench_multi_scene(torch, hiprt, ctx, stream)
        log(f"multi-scene: {extra['multi_scene']}")
        extra["alpha"] = bench_alpha(torch, hiprt, ctx, stream, d_cam, d_diff)
        log(f"alpha-aware tracing: {extra['alpha']}")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as orc
        threads = min(16, os.cpu_count() or 1)
        sample_n = 500_000
        sample = d_diff[:, :sample_n].cpu().numpy()
        orc.time_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, sample[:, :50_000], row=w, threads=threads)
        secs = sorted(orc.time_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, sample, row=w, threads=threads)
                      for _ in range(3))
        one = orc.time_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, sample[:, :50_000], row=w, threads=1)
        cpu = {"value": round(sample_n / secs[1] / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"first {sample_n} first-bounce cosine rays of the 1920x1080 frame, CPU restatement of "
                         f"binary_bvh_tracer (same BVH), OpenMP over rows, median of 3; 1 thread on 50k: "
                         f"{50_000 / one / 1e6:.2f} Mrays/s"}
        log(f"cpu baseline: {cpu}")

    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # HBM bytes per launch of the frame's trace kernel, from the committed
    # FETCH_SIZE / WRITE_SIZE counter summary (null when it holds no entry
    # for this instantiation)
    kname = kernel_name(frame_kw, ds_frame, ds_frame.bvh_format == hiprt.BVH2)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "trace_hall_traffic.json")
    if os.path.exists(tpath):
        entry = json.load(open(tpath)).get("kernels", {}).get(kname.split(" (")[0])
        traffic = entry["bytes_per_launch"] if entry else None

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_b, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hall 262,144 triangles, 1920x1080, 4-bounce cosine path tracing, "
                               "RR p=0.8 at bounces 3-4, batched persistent+graph, "
                               f"{samples} sample(s) per pixel per step",
                   "triangles": int(len(idx)), "width": w, "height": h, "max_depth": 4,
                   "bvh": "SAH 16 planes, 4 tris/leaf, ESC",
                   "layout": "bvh2" if ds_frame.bvh_format == hiprt.BVH2 else "bvh4",
                   "samples_per_step": samples,
                   "rays_per_step": int(round(rays_b))},
        "plain_value": round(plain_value, 1),
        "plain_ms_per_step": round(ms_p, 4),
        "batched_over_plain": round(value / plain_value, 3),
        "trace_kernels_mrays": kernels,
        "autotuned": tuned,
        "frame_config_trial_mrays": {k: round(v, 1) for k, v in trial.items()},
        "one_sample_in_flight_mrays": round(one_sample, 1),
        "frame_config": dict(frame_kw, layout="bvh2" if ds_frame.bvh_format == hiprt.BVH2 else "bvh4")
                        if frame_kw else "library defaults",

        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname,
                     "avg_launch_ms": round(float(ms_lvl.sum()) / max(launches, 1), 4),
                     "bytes_per_launch": round(float(bytes_lvl.sum()) / max(launches, 1)),
                     "compulsory_GBs": round(compulsory, 1)},
        "cpu_baseline": cpu,
        "adjoint_closest": adjoint,
        "refit": refit,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    batched.close()
    plain.close()
    ds4.close()
    ds_tuned.close()
    ds.close()
    ctx.close()


if __name__ == "__main__":
    main()
