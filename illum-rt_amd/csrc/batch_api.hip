// C ABI of the multi-scene batched trace and of the autotuner.
#include "capi_internal.h"

#include "device/kernels.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <vector>

using hiprt_detail::dbuf;
using hiprt_detail::fail;
using hiprt_detail::ok;

namespace {

bool full_rays(const hiprt_rays &r) {
    return r.d_ox && r.d_oy && r.d_oz && r.d_tmin && r.d_dx && r.d_dy && r.d_dz && r.d_tmax;
}
bool full_hits(const hiprt_hits &h) { return h.d_t && h.d_u && h.d_v && h.d_id; }

rt::dev::persistent_cfg cfg_for(int tracer, const hiprt_tracer_opts *opts) {
    rt::dev::persistent_cfg c{};
    c.speculative = tracer >= HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE;
    c.refill = tracer == HIPRT_TRACER_DYNAMIC_WHILE_WHILE;
    c.refill_thr = (opts && opts->refill_threshold > 0) ? std::min(opts->refill_threshold, 64) : 40;
    c.blocks = (opts && opts->blocks > 0) ? opts->blocks : 0;
    c.stats = nullptr;
    c.stack = (opts && opts->stack_config > 0 && opts->stack_config < rt::dev::kStackConfigs) ? opts->stack_config : 0;
    c.chunks = (opts && opts->work_chunks == 1) ? 1 : 8;
    c.park = (opts && opts->park_leaves > 0) ? std::min(opts->park_leaves, rt::dev::kParkClosest)
                                             : rt::dev::kParkClosest;
    return c;
}

} // namespace

extern "C" {

int hiprt_trace_batch(const hiprt_batch_entry *entries, uint32_t n_entries, int32_t any, int32_t tracer,
                      const hiprt_tracer_opts *opts, void *stream) {
    if (!entries && n_entries) return fail(HIPRT_ERR_ARG, "null entries");
    const bool alpha = (tracer & HIPRT_TRACER_ALPHA) != 0;
    tracer &= ~HIPRT_TRACER_ALPHA;
    if (tracer < HIPRT_TRACER_PERSISTENT_WHILE_WHILE || tracer >= HIPRT_TRACER_COUNT)
        return fail(HIPRT_ERR_ARG, "the batched trace needs a persistent tracer");
    if (n_entries == 0) return ok();
    hiprt_context ctx = nullptr;
    // group entries by (BVH format, triangle format)
    std::map<std::pair<int, int>, std::vector<rt::dev::batch_desc>> groups;
    std::map<std::pair<int, int>, uint64_t> totals;
    std::map<std::pair<int, int>, int> need;
    for (uint32_t k = 0; k < n_entries; ++k) {
        const hiprt_batch_entry &e = entries[k];
        if (!e.scene) return fail(HIPRT_ERR_ARG, "entry " + std::to_string(k) + ": null scene");
        if (!ctx) ctx = e.scene->ctx;
        if (e.scene->ctx != ctx) return fail(HIPRT_ERR_ARG, "entries from different contexts");
        if (e.n == 0) continue;
        if (!full_rays(e.rays)) return fail(HIPRT_ERR_ARG, "entry " + std::to_string(k) + ": null ray plane");
        if (!any && !full_hits(e.hits)) return fail(HIPRT_ERR_ARG, "entry " + std::to_string(k) + ": null hit plane");
        if (any && !e.d_occluded && !full_hits(e.hits))
            return fail(HIPRT_ERR_ARG, "entry " + std::to_string(k) + ": any-hit needs occlusion or hit planes");
        auto key = std::make_pair((int)e.scene->desc.bvh, (int)e.scene->desc.tri);
        rt::dev::batch_desc d{};
        d.scene = hiprt_detail::desc_for(e.scene, alpha);
        d.rays = rt::ray_soa{e.rays.d_ox, e.rays.d_oy, e.rays.d_oz, e.rays.d_tmin,
                             e.rays.d_dx, e.rays.d_dy, e.rays.d_dz, e.rays.d_tmax};
        d.hits = full_hits(e.hits) ? rt::hit_soa{e.hits.d_t, e.hits.d_u, e.hits.d_v, e.hits.d_id}
                                   : rt::hit_soa{nullptr, nullptr, nullptr, nullptr};
        d.occluded = e.d_occluded;
        d.first = (uint32_t)totals[key];
        d.n = e.n;
        totals[key] += e.n;
        if (totals[key] > 0xF0000000ull) return fail(HIPRT_ERR_ARG, "too many rays in one batch");
        need[key] = std::max(need[key], e.scene->stack_need);
        groups[key].push_back(d);
    }
    if (!ctx) return ok();
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the table buffer may still be read by an earlier batch on this stream
    HIP_TRY(hipStreamSynchronize(s));
    size_t bytes = 0;
    for (auto &g : groups) bytes += g.second.size() * sizeof(rt::dev::batch_desc);
    hiprt_stream_state *st = ctx->state(stream);
    if (!st) return fail(HIPRT_ERR_HIP, "cannot allocate the stream's work counters");
    dbuf &table = st->batch_table;
    if (table.bytes < bytes) HIP_TRY(table.alloc(bytes));
    size_t off = 0;
    for (auto &g : groups) {
        int bvh = g.first.first, tri = g.first.second;
        rt::dev::persistent_cfg cfg = cfg_for(tracer, opts);
        rt::scene_desc fmt{};
        fmt.bvh = bvh;
        fmt.tri = tri;
        HIP_TRY(hiprt_detail::attach_spill(st, s, fmt, any != 0, true, need[g.first], cfg, opts ? opts->spill_limit : 0));
        auto *dst = reinterpret_cast<rt::dev::batch_desc *>(static_cast<char *>(table.p) + off);
        HIP_TRY(hipMemcpy(dst, g.second.data(), g.second.size() * sizeof(rt::dev::batch_desc), hipMemcpyHostToDevice));
        HIP_TRY(rt::dev::launch_trace_batch(dst, (uint32_t)g.second.size(), (uint32_t)totals[g.first], tri, bvh,
                                            any != 0, st->counter.as<uint32_t>(), cfg, s));
        off += g.second.size() * sizeof(rt::dev::batch_desc);
    }
    return ok();
}

int hiprt_autotune(hiprt_context ctx, hiprt_host_scene s, hiprt_bvh b, const hiprt_rays *sample, uint32_t n,
                   int32_t any, hiprt_tune_result *result, hiprt_scene *tuned) {
    if (!ctx || !s || !b || !sample || !result || !tuned) return fail(HIPRT_ERR_ARG, "null argument");
    if (!full_rays(*sample)) return fail(HIPRT_ERR_ARG, "null ray plane");
    if (n == 0) return fail(HIPRT_ERR_ARG, "empty sample");
    HIP_TRY(hipSetDevice(ctx->device));
    // the two committed layouts and the timing events are released on every
    // return path; the winner's layout is handed to the caller
    struct owned {
        hiprt_scene layouts[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~owned() {
            for (hiprt_scene l : layouts)
                if (l) hiprt_scene_destroy(l);
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } own;
    hiprt_scene *layouts = own.layouts;
    for (int f = 0; f < 2; ++f) {
        int st = hiprt_scene_commit(ctx, s, b, f, HIPRT_TRI_EXACT, &layouts[f]);
        if (st != HIPRT_OK) {
            return st;
        }
    }
    // A persistent grid holds ~300k lanes on MI355X: a batch much smaller than
    // that times tail latency, not throughput, and ranks configurations
    // wrongly.  Small samples are therefore tiled up to kMinTuneRays rays;
    // results are checked on the first n (the sample itself).
    constexpr uint32_t kMinTuneRays = 1u << 20;
    const uint32_t reps = n >= kMinTuneRays ? 1u : (kMinTuneRays + n - 1) / n;
    const uint64_t nt64 = (uint64_t)n * reps;
    if (nt64 > 0xF0000000ull) {
        return fail(HIPRT_ERR_ARG, "sample too large");
    }
    const uint32_t nt = (uint32_t)nt64;
    dbuf tiled;
    hiprt_rays tr = *sample;
    if (reps > 1) {
        if (tiled.alloc(32ull * nt) != hipSuccess) {
            return fail(HIPRT_ERR_HIP, "cannot allocate the tiled sample");
        }
        const float *src[8] = {sample->d_ox, sample->d_oy, sample->d_oz, sample->d_tmin,
                               sample->d_dx, sample->d_dy, sample->d_dz, sample->d_tmax};
        const float *dst[8];
        for (int k = 0; k < 8; ++k) {
            float *plane = tiled.as<float>() + (uint64_t)k * nt;
            dst[k] = plane;
            for (uint32_t r = 0; r < reps; ++r)
                HIP_TRY(hipMemcpy(plane + (uint64_t)r * n, src[k], 4ull * n, hipMemcpyDeviceToDevice));
        }
        tr = hiprt_rays{dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dst[6], dst[7]};
    }
    // reference hits of the sample: plain kernel, binary layout
    dbuf hb, ref_hb, occ;
    if (hb.alloc(16ull * nt) != hipSuccess || ref_hb.alloc(16ull * n) != hipSuccess || occ.alloc(nt) != hipSuccess) {
        return fail(HIPRT_ERR_HIP, "cannot allocate the sample's hit buffers");
    }
    auto hits_of = [](dbuf &d, uint32_t m) {
        float *p = d.as<float>();
        return hiprt_hits{p, p + m, p + 2ull * m, reinterpret_cast<uint32_t *>(p + 3ull * m)};
    };
    hiprt_hits ref_h = hits_of(ref_hb, n), h = hits_of(hb, nt);
    auto trace = [&](hiprt_scene sc, int tracer, const hiprt_tracer_opts *o, hiprt_hits *hh, uint32_t m) {
        return any ? hiprt_trace_any(sc, &tr, m, occ.as<uint8_t>(), hh, tracer, o, nullptr)
                   : hiprt_trace_closest(sc, &tr, m, hh, tracer, o, nullptr);
    };
    if (int st = trace(layouts[0], HIPRT_TRACER_WHILE_WHILE, nullptr, &ref_h, n)) return st;
    std::vector<float> ref_t(n), got_t(n);
    std::vector<uint32_t> ref_id(n), got_id(n);
    if (hipMemcpy(ref_t.data(), ref_h.d_t, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ref_id.data(), ref_h.d_id, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess) {
        return fail(HIPRT_ERR_HIP, "cannot read the reference hits");
    }
    HIP_TRY(hipEventCreate(&own.ev[0]));
    HIP_TRY(hipEventCreate(&own.ev[1]));
    hipEvent_t e0 = own.ev[0], e1 = own.ev[1];
    hiprt_tune_result best{};
    float best_ms = 1e30f;
    int count = 0;
    for (int f = 0; f < 2; ++f)
    for (int stack = 0; stack < rt::dev::kStackConfigs; ++stack) {
        for (int tracer = HIPRT_TRACER_PERSISTENT_WHILE_WHILE; tracer < HIPRT_TRACER_COUNT; ++tracer) {
            std::vector<int> thresholds = tracer == HIPRT_TRACER_DYNAMIC_WHILE_WHILE ? std::vector<int>{24, 40, 56}
                                                                                     : std::vector<int>{40};
            hiprt_tracer_opts base{40, 0, nullptr, stack, 1, 0, 0};
            rt::dev::persistent_cfg pc = cfg_for(tracer, &base);
            int full = 0;
            HIP_TRY(rt::dev::persistent_grid(layouts[f]->desc, any != 0, pc, &full));
            for (int thr : thresholds)
                for (int chunks : {1, 8}) {
                    hiprt_tracer_opts o{thr, full, nullptr, stack, chunks, 0, 0};
                    if (int st = trace(layouts[f], tracer, &o, &h, nt)) return st;
                    float ms = 1e30f;
                    for (int rep = 0; rep < 2; ++rep) {
                        HIP_TRY(hipEventRecord(e0, nullptr));
                        (void)trace(layouts[f], tracer, &o, &h, nt);
                        HIP_TRY(hipEventRecord(e1, nullptr));
                        HIP_TRY(hipEventSynchronize(e1));
                        float m = 0;
                        HIP_TRY(hipEventElapsedTime(&m, e0, e1));
                        ms = std::min(ms, m);
                    }
                    ++count;
                    // results must not depend on the configuration
                    HIP_TRY(hipMemcpy(got_t.data(), h.d_t, 4ull * n, hipMemcpyDeviceToHost));
                    HIP_TRY(hipMemcpy(got_id.data(), h.d_id, 4ull * n, hipMemcpyDeviceToHost));
                    bool same = true;
                    for (uint32_t i = 0; i < n && same; ++i)
                        if (got_id[i] != ref_id[i] && std::memcmp(&got_t[i], &ref_t[i], 4) != 0) same = false;
                    if (!same) continue;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = hiprt_tune_result{tracer, thr, o.blocks, f, (float)(nt / (ms * 1e3)), 0, stack, chunks,
                                                 rt::dev::kParkClosest};
                    }
                }
        }
    }
    if (best_ms >= 1e30f) return fail(HIPRT_ERR_UNSUPPORTED, "no configuration reproduced the plain kernel's hits");
    // the postponement threshold of a speculative winner: parked leaves per lane
    if (!any && best.tracer >= HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE) {
        for (int park = 1; park < rt::dev::kParkClosest; ++park) {
            hiprt_tracer_opts o{best.refill_threshold, best.blocks, nullptr, best.stack_config, best.work_chunks, 0,
                                park};
            hiprt_scene lay = layouts[best.bvh_format];
            if (int st = trace(lay, best.tracer, &o, &h, nt)) return st;
            HIP_TRY(hipEventRecord(e0, nullptr));
            (void)trace(lay, best.tracer, &o, &h, nt);
            HIP_TRY(hipEventRecord(e1, nullptr));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            ++count;
            HIP_TRY(hipMemcpy(got_t.data(), h.d_t, 4ull * n, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(got_id.data(), h.d_id, 4ull * n, hipMemcpyDeviceToHost));
            bool same = true;
            for (uint32_t i = 0; i < n && same; ++i)
                if (got_id[i] != ref_id[i] && std::memcmp(&got_t[i], &ref_t[i], 4) != 0) same = false;
            if (same && ms < best_ms) {
                best_ms = ms;
                best.park_leaves = park;
                best.mrays_per_s = (float)(nt / (ms * 1e3));
            }
        }
    }
    best.candidates = count;
    *result = best;
    *tuned = layouts[best.bvh_format];
    layouts[best.bvh_format] = nullptr;
    return ok();
}

} // extern "C"
