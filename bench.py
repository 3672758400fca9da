#!/usr/bin/env python3
"""Benchmark: closest-hit Mrays/s of the 4-bounce diffuse path tracer on the
262,144-triangle hall at 1920x1080 -- batched (persistent kernel, compacted
queues, hipGraph replay) against plain per-bounce launches.

One step = one path-traced frame: camera rays plus up to four cosine-weighted
diffuse bounces (Russian roulette p = 0.8 at bounces 3 and 4), every level's
rays traced by the gfx950 closest-hit kernel.  `value` counts every traced
ray of the timed frames over their wall time on the GPU stream (inputs -- the
flattened BVH and triangle records -- already resident in HBM).

Prints one JSON line.  Also reported: the plain schedule's rate and the
ratio at equal samples per launch, rates per path level (primary, diffuse),
the trace kernels alone on primary and on diffuse rays, the roofline of the
dominant kernel and the CPU oracle's throughput on the same diffuse rays.

Roofline: the HBM-side bytes of the dominant kernel (the persistent
closest-hit trace of the timed frame configuration) are measured in this run
by two rocprofv3 counter passes over a child process that renders the same
configuration -- FETCH_SIZE, then WRITE_SIZE, each scaled by what the counter
reads for a 1 GiB device copy in the same child -- and divided by the
kernel's average launch time from HIP events on its stream.  A third pass
reads the VALU counters (instructions, active lanes), the roof that binds,
and a kernel-trace pass times the same dispatches as rocprofv3 sees them.
The CSVs stay under COUNTER_DIR (one folder per pass).
"""
import argparse
import csv
import glob
import json
import os
import shutil
import signal
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
PEAK_CLOCK_GHZ = 2.4   # MI355X peak engine clock
SIMDS = 256 * 4        # CUs x SIMDs per CU
METRIC = "Mrays/s closest-hit, 262k-tri scene, diffuse rays: batched vs plain kernel"
COUNTER_DIR = os.path.join(ROOT, "gpurun_out", "bench_counters")


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


def rocprof_bin():
    for p in (shutil.which("rocprofv3"), "/opt/rocm/bin/rocprofv3"):
        if p and os.path.exists(p):
            return p
    return None


def read_counters(outdir):
    """Per-dispatch counter values from the rocprofv3 counter-collection CSVs
    under outdir, in dispatch order: [{"kernel": name, counter: value, ...}]
    (values of one counter in one dispatch summed over its rows)."""
    rows = {}
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                key = (r.get("Process_Id", ""), int(r.get("Dispatch_Id") or r.get("Correlation_Id") or 0))
                e = rows.setdefault(key, {"kernel": r.get("Kernel_Name", "")})
                name = r.get("Counter_Name", "")
                e[name] = e.get(name, 0.0) + float(r.get("Counter_Value") or 0.0)
    return [rows[k] for k in sorted(rows, key=lambda k: k[1])]


def profiled_child(rp_args, child_cfg, tag, timeout=300):
    """Runs this script as a counter child (counter_child) under
    `rocprofv3 <rp_args>`, output CSVs in COUNTER_DIR/tag; returns that folder
    and rocprofv3's exit code."""
    rp = rocprof_bin()
    if rp is None:
        raise RuntimeError("rocprofv3 not found")
    outdir = os.path.join(COUNTER_DIR, tag)
    shutil.rmtree(outdir, ignore_errors=True)
    os.makedirs(outdir, exist_ok=True)
    cmd = [rp, *rp_args, "--output-format", "csv", "-d", outdir, "-o", "pass", "--",
           sys.executable, os.path.abspath(__file__), "--counter-child", json.dumps(child_cfg)]
    with open(os.path.join(outdir, "child.log"), "w") as logf:
        proc = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=logf,
                                stderr=subprocess.STDOUT, start_new_session=True)
        try:
            rc = proc.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            raise RuntimeError(f"rocprofv3 pass {tag} timed out")
    return outdir, rc


def counter_pass(counters, child_cfg, tag, timeout=300):
    """Per-dispatch counters (read_counters) of a counter child run under
    `rocprofv3 --pmc <counters>`."""
    outdir, rc = profiled_child(["--pmc", *counters], child_cfg, tag, timeout)
    rows = read_counters(outdir)
    if not rows:
        raise RuntimeError(f"counter pass {tag}: no counter rows (rocprofv3 exit {rc})")
    return rows


def trace_pass(child_cfg, kname, skip, tag="trace", timeout=300):
    """Durations (ms) of the dispatches of kernel `kname` after the first
    `skip`, from a counter child run under `rocprofv3 --kernel-trace --stats`."""
    outdir, rc = profiled_child(["--kernel-trace", "--stats"], child_cfg, tag, timeout)
    durs = []
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*kernel_trace.csv"), recursive=True)):
        with open(f, newline="") as fh:
            rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Dispatch_Id"]))
        durs += [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6 for r in rows
                 if kname in r["Kernel_Name"]]
    if len(durs) <= skip:
        raise RuntimeError(f"kernel-trace pass: no dispatches of {kname} (rocprofv3 exit {rc})")
    return durs[skip:]


def counter_child(cfg):
    """--counter-child: render the frame configuration `cfg` (one warm-up
    frame, then cfg["frames"]), then copy 1 GiB device to device (the
    calibration dispatch)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "illum-rt_amd"))
    import hiprt
    ctx = hiprt.Context(0)
    sc = hiprt.HostScene.generate(3)
    bvh = hiprt.Bvh(sc)
    ds = hiprt.DeviceScene(ctx, sc, bvh, hiprt.BVH4 if cfg["layout"] == "bvh4" else hiprt.BVH2)
    pt = hiprt.PathTracer(ds, sc.cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True, **cfg["kw"])
    for f in range(1 + cfg["frames"]):
        pt.render(f)
    torch.cuda.synchronize()
    pt.close()
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    src.fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    ds.close()
    ctx.close()


def frame_dispatches(rows, kname, counter, skip):
    """Values of `counter` in the dispatches of kernel `kname` after the first `skip`."""
    vals = [r[counter] for r in rows if kname in r["kernel"] and counter in r]
    return vals[skip:]


def copy_factor(rows, counter):
    """Bytes per counter unit: 1 GiB over what `counter` reads (in KiB) for the
    child's 1 GiB device copy."""
    vals = [r[counter] for r in rows if "copyBuffer" in r["kernel"] and counter in r]
    if not vals or vals[-1] <= 0:
        raise RuntimeError(f"no 1 GiB copy dispatch with {counter}")
    return float(1 << 30) / vals[-1], vals[-1]


def measure_roofline(child_cfg, kname, levels, avg_launch_ms):
    """HBM bytes per launch of the frame kernel and its VALU figures, from three
    counter passes, and its average launch time under the kernel trace (to
    set beside the HIP-event time); dict for the JSON line."""
    out = {}
    fetch = counter_pass(["FETCH_SIZE"], child_cfg, "fetch")
    write = counter_pass(["WRITE_SIZE"], child_cfg, "write")
    ff, f_copy = copy_factor(fetch, "FETCH_SIZE")
    wf, w_copy = copy_factor(write, "WRITE_SIZE")
    f = frame_dispatches(fetch, kname, "FETCH_SIZE", levels)
    w = frame_dispatches(write, kname, "WRITE_SIZE", levels)
    if not f or not w:
        raise RuntimeError(f"no dispatches of {kname} in the counter passes")
    read_b = sum(f) / len(f) * ff
    write_b = sum(w) / len(w) * wf
    out["traffic"] = round(read_b + write_b)
    out["traffic_read"] = round(read_b)
    out["traffic_write"] = round(write_b)
    out["traffic_dispatches"] = len(f)
    out["calibration"] = {"copy_bytes": 1 << 30, "fetch_kib": round(f_copy, 1), "write_kib": round(w_copy, 1),
                          "fetch_bytes_per_unit": round(ff, 2), "write_bytes_per_unit": round(wf, 2)}
    valu = counter_pass(["SQ_INSTS_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_ACTIVE_INST_VALU", "GRBM_GUI_ACTIVE"],
                        child_cfg, "valu")
    insts = frame_dispatches(valu, kname, "SQ_INSTS_VALU", levels)
    thr = frame_dispatches(valu, kname, "SQ_THREAD_CYCLES_VALU", levels)
    act = frame_dispatches(valu, kname, "SQ_ACTIVE_INST_VALU", levels)
    if insts and thr and act:
        i_mean = sum(insts) / len(insts)
        cycles = avg_launch_ms * 1e-3 * PEAK_CLOCK_GHZ * 1e9
        out["valu"] = {"insts_per_launch": round(i_mean),
                       "busy_frac": round(i_mean * 4.0 / (SIMDS * cycles), 4),
                       "active_lanes": round(sum(thr) / max(sum(act), 1.0), 2),
                       "active_lane_frac": round(sum(thr) / max(sum(act), 1.0) / 64.0, 4),
                       "basis": "SQ_INSTS_VALU x 4 cycles per wave64 instruction over 1024 SIMDs x launch time x "
                                "2.4 GHz (an issue-rate estimate: it passes 1 when some counted instructions "
                                "take fewer than 4 cycles of the SIMD or issue beside it, as transcendental "
                                "ones do); active lanes = SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU"}
    durs = trace_pass(child_cfg, kname, levels)
    out["kernel_trace"] = {"avg_launch_ms": round(sum(durs) / len(durs), 4), "dispatches": len(durs),
                           "basis": "rocprofv3 --kernel-trace over the counter child's frames, warm-up frame "
                                    "excluded"}
    return out


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


def bench_adjoint(torch, hiprt, ds, d_diff, stream, reps=10):
    """Closest-hit adjoint on 1M of the frame's first-bounce diffuse rays:
    vertex, origin and direction gradients of sum(a t + b beta + c gamma)
    with the hit set frozen; the vertex gradients are summed per hit
    triangle after a sort."""
    n_adj = 1_000_000
    adj_rays = d_diff[:, :n_adj].contiguous()
    adj_hits = hiprt.new_hits(n_adj)
    hiprt.trace_closest(ds, adj_rays, adj_hits, tracer="persistent-while-while", stream=stream.cuda_stream)
    gen = torch.Generator(device="cuda").manual_seed(5)
    g_t, g_u, g_v = (torch.randn(n_adj, device="cuda", generator=gen) for _ in range(3))
    g_vert = torch.zeros((hiprt.scene_counts(ds)[0], 3), dtype=torch.float32, device="cuda")
    g_o = torch.empty((n_adj, 3), dtype=torch.float32, device="cuda")
    g_d = torch.empty((n_adj, 3), dtype=torch.float32, device="cuda")
    ms_adj = time_calls(torch, lambda: hiprt.adjoint_closest(ds, adj_rays, adj_hits, g_t, g_u, g_v, g_vert, g_o, g_d,
                                                             stream=stream.cuda_stream), stream, reps=reps)
    # algorithmic bytes per ray: the per-ray pass reads the ray (24), hit (8),
    # upstream gradients (12) and triangle (48) and writes origin and
    # direction gradients (24), a corner record (36) and a key (4); the sort
    # moves key and index 3 times (3 x 16); the vertex pass reads key, index
    # and corner record (44)
    per_ray = 24 + 8 + 12 + 48 + 24 + 36 + 4 + 48 + 44
    return {"rays": n_adj, "ms": round(ms_adj, 4), "mrays": round(n_adj / ms_adj / 1e3, 1),
            "bytes_per_ray": per_ray, "achieved_GBs": round(n_adj * per_ray / (ms_adj * 1e-3) / 1e9, 1)}


def bench_regen(torch, hiprt, ds, cam, kw, stream, total=64, frames=2):
    """Path regeneration against samples in flight at equal memory: `total`
    samples per pixel rendered by a pool of k path slots per pixel that
    refill themselves (one render call), and by k samples in flight
    (total / k render calls), for k = 1, 4, 16; frame configuration `kw`."""
    base = {k: v for k, v in kw.items() if k != "samples"}
    out = {"samples_per_pixel": total, "config": base}

    def run(pt, calls):
        f = 0
        for _ in range(calls):   # warm-up
            pt.render(f, stream.cuda_stream)
            f += 1
        torch.cuda.synchronize()
        before = pt.ray_totals()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(frames * calls):
            pt.render(f, stream.cuda_stream)
            f += 1
        e1.record(stream)
        torch.cuda.synchronize()
        rays = float((pt.ray_totals() - before).sum())
        return e0.elapsed_time(e1) / frames, rays / frames

    def one(extra, calls):
        pt = hiprt.PathTracer(ds, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True,
                              **base, **extra)
        ms, rays = run(pt, calls)
        mem = pt.memory()
        pt.close()
        return {"ms_per_frame": round(ms, 3), "mrays": round(rays / ms / 1e3, 1),
                "device_mib": round(mem / 2 ** 20, 1)}, mem

    for k in (1, 4, 16):
        row = {}
        row["in_flight"], m_if = one(dict(samples=k), total // k)
        row["regen"], m_r = one(dict(samples=total, regen_pool=k), 1)
        # a regeneration slot holds fewer bytes: the pool that fits the
        # in-flight schedule's memory
        k_eq = int(k * m_if / m_r)
        if k_eq > k:
            row[f"regen_pool_{k_eq}_equal_memory"], _ = one(dict(samples=total, regen_pool=k_eq), 1)
        out[f"pool_{k}"] = row
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos, idx, bvh, d_diff, row):
    """The oracle's restatement of the reference's CPU tracer on the frame's
    2,073,600 first-bounce diffuse rays (the GPU's buffer, same BVH), OpenMP
    over rows of the image: 1 warm-up + median of 5 on the threads of this
    process's CPU share, and the median of 3 on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    affinity = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(omp, affinity) if omp > 0 else affinity
    rays = d_diff.cpu().numpy()
    n = rays.shape[1]
    run = lambda t: orc.time_closest(pos, idx, bvh.nodes, bvh.root, bvh.index, rays, row=row, threads=t)
    run(threads)
    secs = sorted(run(threads) for _ in range(5))
    one = sorted(run(1) for _ in range(3))
    return {"value": round(n / secs[2] / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"all {n} first-bounce cosine rays of the 1920x1080 frame (the GPU's buffer), CPU restatement "
                      "of binary_bvh_tracer::closest_hit over the same BVH, OpenMP over image rows "
                      "(schedule static), 1 warm-up + median of 5",
            "one_thread": round(n / one[1] / 1e6, 3),
            "host": {"model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": affinity,
                     "omp_num_threads": omp or None}}


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
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the closest-hit adjoint on 1M of the frame's diffuse rays (for profiling)")
    ap.add_argument("--regen-only", action="store_true",
                    help="only compare path regeneration with samples in flight (stack configuration 4, "
                         "dynamic-while-while, binary tree)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (roofline traffic and VALU figures)")
    ap.add_argument("--counter-child", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.counter_child:
        counter_child(json.loads(args.counter_child))
        return
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

    if args.adjoint_only:
        print(json.dumps(bench_adjoint(torch, hiprt, ds, d_diff, stream, reps=20)), flush=True)
        ds.close()
        ctx.close()
        return

    if args.regen_only:
        print(json.dumps(bench_regen(torch, hiprt, ds, cam, {"tracer": "dynamic-while-while", "stack_config": 4},
                                     stream)), flush=True)
        ds.close()
        ctx.close()
        return

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

    # The frame configuration comes from the library's frame-level tuner
    # (hiprt_pt_autotune): trial frames of each candidate -- tracer, layout,
    # stack configuration at one sample per pixel, then samples in flight up
    # to 64, then the parked-leaf limit -- in rays per second.  Its one-sample
    # winner is the like-for-like batched configuration.
    ds4 = hiprt.DeviceScene(ctx, sc, bvh, hiprt.BVH4)
    t_tune = time.time()
    frame_kw, frame_tune = hiprt.pt_autotune(ds, ds4, cam, max_samples=64)
    frame_tune["seconds"] = round(time.time() - t_tune, 1)
    one_kw, one_tune = hiprt.pt_autotune(ds, ds4, cam, max_samples=1)
    log(f"frame tuner: {frame_kw} {frame_tune}; one sample: {one_kw} {one_tune}")
    ds_frame = ds4 if frame_tune["layout"] == "bvh4" else ds
    ds_one = ds4 if one_tune["layout"] == "bvh4" else ds
    samples = frame_kw.get("samples", 1)

    batched = hiprt.PathTracer(ds_frame, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True,
                               **frame_kw)
    plain = hiprt.PathTracer(ds, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=False)
    ms_b, rays_b, per_level_b = time_frames(torch, batched, args.steps, args.warmup, stream)
    log(f"batched: {ms_b:.3f} ms/frame, {rays_b / 1e6:.3f} Mrays/frame, per level {per_level_b.tolist()}")
    ms_p, rays_p, per_level_p = time_frames(torch, plain, args.steps, args.warmup, stream)
    log(f"plain:   {ms_p:.3f} ms/frame, {rays_p / 1e6:.3f} Mrays/frame")
    # like for like: the batched schedule with one sample per pixel per
    # launch, as the plain schedule traces
    one = hiprt.PathTracer(ds_one, cam, max_depth=4, rr_start=3, rr_survive=0.8, seed=1, batched=True, **one_kw)
    ms_1, rays_1, _ = time_frames(torch, one, args.steps, args.warmup, stream)
    one.close()
    log(f"batched, one sample per launch ({one_kw}): {ms_1:.3f} ms/frame, {rays_1 / 1e6:.3f} Mrays/frame")
    value = rays_b / ms_b / 1e3
    plain_value = rays_p / ms_p / 1e3
    one_value = rays_1 / ms_1 / 1e3
    # Dominant kernel: the persistent closest-hit trace.  Per level: HIP-event
    # time of the launch, and the counting variant's fetches on the same rays.
    node_bytes = 64.0 if (ds_frame.bvh_format == hiprt.BVH2) else 128.0
    ms_lvl, fetch, rays_lvl = batched.profile(frame=10_000, stream=stream.cuda_stream)
    fetched_lvl = 48.0 * rays_lvl + node_bytes * fetch[:, 0] + 48.0 * fetch[:, 1]
    launches = int((rays_lvl > 0).sum())
    avg_launch_ms = float(ms_lvl.sum()) / max(launches, 1)
    footprint = ds_frame.n_nodes * node_bytes + ds_frame.n_slots * 48
    compulsory_lvl = 48.0 * rays_lvl + footprint * (rays_lvl > 0)
    log(f"trace per level ms {np.round(ms_lvl, 3).tolist()}, nodes/ray "
        f"{np.round(fetch[:, 0] / np.maximum(rays_lvl, 1), 1).tolist()}, tris/ray "
        f"{np.round(fetch[:, 1] / np.maximum(rays_lvl, 1), 1).tolist()}")
    per_level = {"rays": [int(x) for x in rays_lvl], "trace_ms": [round(float(x), 4) for x in ms_lvl],
                 "trace_mrays": [round(float(r) / float(m) / 1e3, 1) if m > 0 else None
                                 for r, m in zip(rays_lvl, ms_lvl)]}
    diffuse_trace = float(rays_lvl[1:].sum()) / float(ms_lvl[1:].sum()) / 1e3
    primary_trace = float(rays_lvl[0]) / float(ms_lvl[0]) / 1e3
    diffuse_per_step = float(per_level_b[1:].sum()) / ms_b / 1e3

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

    regen = bench_regen(torch, hiprt, ds_frame, cam, frame_kw, stream)
    log(f"regeneration vs samples in flight: {regen}")

    adjoint = bench_adjoint(torch, hiprt, ds, d_diff, stream)
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
        cpu = cpu_baseline(pos, idx, bvh, d_diff, w)
        log(f"cpu baseline: {cpu}")

    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # Roofline of the frame's trace kernel: HBM-side bytes per launch from the
    # counter passes, over the kernel's average launch time measured above.
    kname = kernel_name(frame_kw, ds_frame, ds_frame.bvh_format == hiprt.BVH2)
    layout = "bvh2" if ds_frame.bvh_format == hiprt.BVH2 else "bvh4"
    fetched_per_launch = float(fetched_lvl.sum()) / max(launches, 1)
    roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
            "traffic": None, "kernel": kname, "avg_launch_ms": round(avg_launch_ms, 4),
            "fetched_bytes_per_launch": round(fetched_per_launch),
            "fetched_bytes_per_s": round(fetched_per_launch / (avg_launch_ms * 1e-3)),
            "compulsory_bytes_per_launch": round(float(compulsory_lvl.sum()) / max(launches, 1))}
    if not args.no_counters:
        child = {"layout": layout, "kw": frame_kw, "frames": 2}
        try:
            meas = measure_roofline(child, kname.split(" (")[0], 5, avg_launch_ms)
            achieved = meas["traffic"] / (avg_launch_ms * 1e-3) / 1e9
            roof.update(meas)
            roof["achieved"] = round(achieved, 1)
            roof["frac"] = round(achieved / HBM_PEAK_GBS, 4)
        except Exception as e:  # the counters are a measurement, not the product: report why they are missing
            roof["traffic_error"] = str(e)[:300]
        log(f"roofline: {roof}")

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
        # like for like: one sample per pixel per launch on both schedules
        "batched_over_plain": round(one_value / plain_value, 3),
        "one_sample_in_flight_mrays": round(one_value, 1),
        "one_sample_config": dict(one_kw, layout=one_tune["layout"]),
        "batched_over_plain_at_value": round(value / plain_value, 3),
        "diffuse_rays_per_step_mrays": round(diffuse_per_step, 1),
        "diffuse_levels_trace_mrays": round(diffuse_trace, 1),
        "primary_level_trace_mrays": round(primary_trace, 1),
        "per_level": per_level,
        "trace_kernels_mrays": kernels,
        "autotuned": tuned,
        "frame_config": dict(frame_kw, layout=layout),
        "frame_tuner": frame_tune,
        "roofline": roof,
        "cpu_baseline": cpu,
        "regeneration": regen,
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
