// C ABI of the wavefront steps and the wavefront path tracer.
//
// The batched schedule records one frame -- counter resets, camera rays, and
// per level a persistent trace that reads the live-ray count from device
// memory followed by a compacting shade step -- into a hipGraph once, and
// replays it per frame.  Counters are cleared by fill kernels, never by
// memset nodes, and the frame's random seed is set by a one-thread kernel
// launched before the graph, so nothing host-side changes between replays.
#include "capi_internal.h"

#include <algorithm>

#include "core/lights.h"
#include "core/raygen.h"
#include "device/kernels.h"
#include "device/pt_kernels.h"
#include "device/ray_sort.h"
#include "device/steps_kernels.h"
#include "device/adjoint_kernels.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

using hiprt_detail::dbuf;
using hiprt_detail::fail;
using hiprt_detail::ok;

namespace {

// Sort grid over the scene box: the union of the root's child boxes (a
// one-leaf or empty tree sorts by the direction's cell or octant only).
rt::dev::sort_grid grid_of(const rt::flat_scene &h) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto grow = [&](float xl, float xh, float yl, float yh, float zl, float zh) {
        lo[0] = std::min(lo[0], xl);
        lo[1] = std::min(lo[1], yl);
        lo[2] = std::min(lo[2], zl);
        hi[0] = std::max(hi[0], xh);
        hi[1] = std::max(hi[1], yh);
        hi[2] = std::max(hi[2], zh);
    };
    if (h.root_ref >= 0 && h.bvh_kind == rt::BVH2 && (size_t)h.root_ref < h.nodes2.size()) {
        const rt::node2 &r = h.nodes2[h.root_ref];
        grow(r.lxlo, r.lxhi, r.lylo, r.lyhi, r.lzlo, r.lzhi);
        grow(r.rxlo, r.rxhi, r.rylo, r.ryhi, r.rzlo, r.rzhi);
    } else if (h.root_ref >= 0 && h.bvh_kind == rt::BVH4 && (size_t)h.root_ref < h.nodes4.size()) {
        const rt::node4 &r = h.nodes4[h.root_ref];
        for (int k = 0; k < 4; ++k)
            if (r.ref[k] != rt::kInvalidRef) grow(r.lox[k], r.hix[k], r.loy[k], r.hiy[k], r.loz[k], r.hiz[k]);
    }
    rt::dev::sort_grid g{};
    for (int a = 0; a < 3; ++a) {
        float ext = hi[a] - lo[a];
        bool ok_axis = std::isfinite(lo[a]) && std::isfinite(ext) && ext > 0.0f;
        g.lo[a] = ok_axis ? lo[a] : 0.0f;
        g.scale[a] = ok_axis ? (float)(1u << rt::dev::kSortCellBits) / ext : 0.0f;
        if (!std::isfinite(g.scale[a])) g.scale[a] = 0.0f;
    }
    return g;
}

// Shading data of a committed scene for the device steps.
rt::dev::pt_scene pt_scene_of(const hiprt_scene_s *sc) {
    return rt::dev::pt_scene{sc->tri_geo.as<const float4>(), sc->mats.as<const float4>(), sc->n_tris,
                             sc->tri_ns.as<const float4>()};
}

rt::cam_frame frame_from(const float *cam) {
    return rt::make_cam_frame(rt::vec3f{cam[0], cam[1], cam[2]}, rt::vec3f{cam[3], cam[4], cam[5]},
                              rt::vec3f{cam[6], cam[7], cam[8]}, cam[9], (int)cam[10], (int)cam[11]);
}

// Eight planes of n floats carved from one allocation.
struct ray_planes {
    dbuf mem;
    rt::dev::ray_out out{};
    hipError_t alloc(uint32_t n) {
        hipError_t e = mem.alloc(8ull * n * 4);
        if (e != hipSuccess) return e;
        float *b = mem.as<float>();
        out = rt::dev::ray_out{b, b + n, b + 2ull * n, b + 3ull * n, b + 4ull * n, b + 5ull * n, b + 6ull * n,
                               b + 7ull * n};
        return hipSuccess;
    }
};

} // namespace

struct hiprt_pt_s {
    hiprt_scene sc = nullptr;
    rt::cam_frame f{};
    uint32_t n = 0;
    rt::dev::pt_params pp{};
    uint64_t seed = 0;
    int tracer = 0;
    bool alpha = false;   // alpha-aware tracing (tracer id | HIPRT_TRACER_ALPHA)
    bool batched = false;
    bool jitter = true;   // camera rays through a random point of the pixel (else its centre)
    rt::dev::persistent_cfg pcfg{};
    // refill threshold of levels 0 and 1 with a pixel's samples in flight:
    // queued pixel-major, a wavefront holds one pixel's camera rays (and,
    // sorted, rays leaving nearby points), and refilling it early mixes in
    // other pixels' rays
    int refill_coherent = 16;
    ray_planes rays[2];
    dbuf hits, paths[2], thr, rad, counts, totals, frame_seed, work, spill;
    // samples in flight: n = samples * n_px paths; with more than one the
    // paths accumulate radiance in rad_paths, added to `rad` (per pixel) at the
    // end of the frame in sample order
    uint32_t samples = 1, n_px = 0;
    dbuf rad_paths;
    // sorted bounce queues: ray records (ray, path id, throughput) per level
    // parity, keys, the permutation, and the radix sort's scratch
    bool sorted = false;
    dbuf rec[2], keys, keys_tmp, vals_tmp, perm, sort_counts, sort_totals;
    rt::dev::sort_grid grid{};
    rt::hit_soa h{};
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    bool clear_next = true;
    // path regeneration: `lanes` slots per pixel (0: off); queue lengths of
    // the two record buffers, per-slot radiance (lanes > 1), the graph of
    // kRegenChunk iterations, and the host's view of the queue length
    uint32_t lanes = 0;
    dbuf qcount, acc;
    hipGraphExec_t regen_exec = nullptr;
    uint32_t *host_q = nullptr;
    hipEvent_t regen_ev[2] = {nullptr, nullptr};
    static constexpr int kRegenChunk = 8;   // even: every chunk starts on record buffer 0

    ~hiprt_pt_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (regen_exec) (void)hipGraphExecDestroy(regen_exec);
        for (hipEvent_t e : regen_ev)
            if (e) (void)hipEventDestroy(e);
        if (host_q) (void)hipHostFree(host_q);
        if (cap) (void)hipStreamDestroy(cap);
    }

    rt::dev::regen_params regen() const {
        return rt::dev::regen_params{f,     frame_seed.as<const uint64_t>(), jitter ? 1 : 0, n_px, samples, lanes,
                                     lanes > 1 ? acc.as<float>() : nullptr, rad.as<float>(), counts.as<uint32_t>()};
    }

    rt::dev::radix_sort_buffers sort_buffers() const {
        return rt::dev::radix_sort_buffers{keys.as<uint32_t>(),        keys_tmp.as<uint32_t>(),
                                           vals_tmp.as<uint32_t>(),    perm.as<uint32_t>(),
                                           sort_counts.as<uint32_t>(), sort_totals.as<uint32_t>()};
    }

    // kRegenChunk regeneration iterations: trace the queue in record buffer
    // a, shade it into buffer b (continued paths and regenerated camera
    // rays), sort b's keys.
    hipError_t record_regen_chunk(hipStream_t s) {
        using namespace rt::dev;
        uint32_t *q = qcount.as<uint32_t>();
        const regen_params rp = regen();
        for (int it = 0; it < kRegenChunk; ++it) {
            const int a = it & 1, b = a ^ 1;
            ray_gather g;
            g.aos = rec[a].as<const float4>();
            g.perm = sorted ? perm.as<const uint32_t>() : nullptr;
            g.stride = kRayRecord;
            g.fixed_t = 1;
            hipError_t e = launch_trace_persistent(hiprt_detail::desc_for(sc, alpha), rt::ray_soa{}, h, nullptr, n,
                                                   q + a, work.as<uint32_t>(), false, pcfg, s, g);
            if (e == hipSuccess) e = launch_fill_u32(q + b, 0u, 1u, s);
            if (e == hipSuccess)
                e = launch_shade_regen(ps(), pp, rp, rec[a].as<const float4>(), g.perm, h, rec[b].as<float4>(),
                                       sorted ? keys.as<uint32_t>() : nullptr, q + a, q + b, grid, n, s);
            if (e == hipSuccess && sorted) e = launch_radix_sort(sort_buffers(), n, q + b, regen_key_bits(lanes), s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

    // One render call with path regeneration (frame seeds already set).
    int render_regen(hipStream_t s) {
        using namespace rt::dev;
        uint32_t *q = qcount.as<uint32_t>();
        const int levels = pp.max_depth + 1;
        if (clear_next) HIP_TRY(launch_fill_u32(rad.as<uint32_t>(), 0u, 3u * n_px, s));
        HIP_TRY(launch_fill_u32(counts.as<uint32_t>(), 0u, (uint32_t)levels + 1, s));
        HIP_TRY(launch_fill_u32(q, n, 1u, s));
        HIP_TRY(launch_fill_u32(q + 1, 0u, 1u, s));
        HIP_TRY(launch_regen_init(regen(), rec[0].as<float4>(), sorted ? keys.as<uint32_t>() : nullptr, s));
        if (sorted) HIP_TRY(launch_radix_sort(sort_buffers(), n, q, regen_key_bits(lanes), s));
        // every slot ends within ceil(samples / lanes) paths of at most
        // max_depth + 1 iterations each
        const uint64_t bound = (uint64_t)((samples + lanes - 1) / lanes) * (uint64_t)levels;
        const uint64_t chunks = (bound + kRegenChunk - 1) / kRegenChunk;
        bool drained = false;
        uint64_t c = 0;
        for (; c < chunks && !drained; ++c) {
            HIP_TRY(hipGraphLaunch(regen_exec, s));
            HIP_TRY(hipMemcpyAsync(host_q + (c & 1), q, 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipEventRecord(regen_ev[c & 1], s));
            if (c >= 1) {   // chunk c is queued behind the one waited for
                HIP_TRY(hipEventSynchronize(regen_ev[(c - 1) & 1]));
                drained = host_q[(c - 1) & 1] == 0u;
            }
        }
        if (!drained) {
            HIP_TRY(hipEventSynchronize(regen_ev[(c - 1) & 1]));
            if (host_q[(c - 1) & 1] != 0u)
                return fail(HIPRT_ERR_HIP, "path regeneration did not drain its queue within the iteration bound");
        }
        if (lanes > 1) HIP_TRY(launch_add_samples(rad.as<float>(), acc.as<const float>(), n_px, lanes, s));
        HIP_TRY(launch_add_counts(totals.as<unsigned long long>(), counts.as<const uint32_t>(), levels, s));
        clear_next = false;
        return ok();
    }

    rt::dev::pt_scene ps() const { return pt_scene_of(sc); }

    // The frame's launches on stream s (frame seed already set).  When
    // `ev` is non-null, events ev[2L], ev[2L+1] bracket level L's trace;
    // when `stats` is non-null the counting trace variant adds level L's
    // fetch counts to stats[2L], stats[2L+1].
    hipError_t record(hipStream_t s, bool clear_radiance, hipEvent_t *ev = nullptr,
                      unsigned long long *stats = nullptr) {
        using namespace rt::dev;
        uint32_t *cnt = counts.as<uint32_t>();
        int levels = pp.max_depth + 1;
        hipError_t e = launch_fill_u32(cnt, 0u, (uint32_t)levels + 1, s);
        if (e != hipSuccess) return e;
        e = launch_fill_u32(cnt, n, 1u, s);
        if (e != hipSuccess) return e;
        float *prad = samples > 1 ? rad_paths.as<float>() : rad.as<float>();   // radiance per path
        const uint64_t *fs = frame_seed.as<const uint64_t>();
        // the sorting form's camera level is path i with throughput 1 in slot i:
        // no path ids or throughputs to write
        e = launch_camera_rays(f, rays[0].out, batched && !sorted ? paths[0].as<uint32_t>() : nullptr,
                               sorted ? nullptr : thr.as<float>(),
                               rad.as<float>(), clear_radiance, s, n, samples > 1 ? prad : nullptr,
                               jitter ? fs : nullptr, pp.px_samples);
        if (e != hipSuccess) return e;
        for (int L = 0; L < levels; ++L) {
            if (ev) {
                e = hipEventRecord(ev[2 * L], s);
                if (e != hipSuccess) return e;
            }
            unsigned long long *lstats = stats ? stats + 2 * L : nullptr;
            if (!batched) {
                e = launch_trace_plain(hiprt_detail::desc_for(sc, alpha), rays[0].out.as_in(), h, nullptr, n, false, s,
                                       lstats);
                if (e != hipSuccess) return e;
                if (ev) {
                    e = hipEventRecord(ev[2 * L + 1], s);
                    if (e != hipSuccess) return e;
                }
                e = launch_shade_plain(ps(), pp, L, fs, rays[0].out, h, thr.as<float>(), prad, cnt, n, s);
            } else {
                // persistent tracers read the queue length from counts[L]
                persistent_cfg cfg = pcfg;
                cfg.stats = lstats;
                if (pp.px_samples > 1 && L <= 1) cfg.refill_thr = refill_coherent;
                if (!sorted) {
                    const ray_planes &in = rays[L & 1], &outq = rays[(L + 1) & 1];
                    e = launch_trace_persistent(hiprt_detail::desc_for(sc, alpha), in.out.as_in(), h, nullptr, n,
                                                cnt + L,
                                                work.as<uint32_t>(), false, cfg, s);
                    if (e != hipSuccess) return e;
                    if (ev) {
                        e = hipEventRecord(ev[2 * L + 1], s);
                        if (e != hipSuccess) return e;
                    }
                    e = launch_shade_compact(ps(), pp, L, fs, in.out, h, paths[L & 1].as<const uint32_t>(),
                                             outq.out, paths[(L + 1) & 1].as<uint32_t>(), thr.as<float>(), prad,
                                             cnt, n, s);
                } else {
                    // level 0: camera ray planes in pixel order; later levels:
                    // the previous shade step's records through the sorted
                    // permutation
                    const float4 *in_rec = L > 0 ? rec[(L - 1) & 1].as<const float4>() : nullptr;
                    const uint32_t *pm = L > 0 ? perm.as<const uint32_t>() : nullptr;
                    ray_gather g;
                    g.aos = in_rec;
                    g.perm = pm;
                    g.stride = kRayRecord;
                    e = launch_trace_persistent(hiprt_detail::desc_for(sc, alpha), rays[0].out.as_in(), h, nullptr,
                                                n, cnt + L,
                                                work.as<uint32_t>(), false, cfg, s, g);
                    if (e != hipSuccess) return e;
                    if (ev) {
                        e = hipEventRecord(ev[2 * L + 1], s);
                        if (e != hipSuccess) return e;
                    }
                    e = launch_shade_stage(ps(), pp, L, fs, rays[0].out, nullptr, nullptr, in_rec, pm, h,
                                           rec[L & 1].as<float4>(),
                                           keys.as<uint32_t>(), prad, cnt, grid, n, s);
                    if (e == hipSuccess && L < pp.max_depth) {
                        radix_sort_buffers rb{keys.as<uint32_t>(),      keys_tmp.as<uint32_t>(),
                                              vals_tmp.as<uint32_t>(),  perm.as<uint32_t>(),
                                              sort_counts.as<uint32_t>(), sort_totals.as<uint32_t>()};
                        e = launch_radix_sort(rb, n, cnt + L + 1, pp.px_samples > 1 ? kRayKeyBits : kOctantKeyBits, s);
                    }
                }
            }
            if (e != hipSuccess) return e;
        }
        if (samples > 1) {
            e = launch_add_samples(rad.as<float>(), rad_paths.as<const float>(), n_px, samples, s,
                                   pp.px_samples > 1);
            if (e != hipSuccess) return e;
        }
        return launch_add_counts(totals.as<unsigned long long>(), cnt, levels, s);
    }
};

extern "C" {

int hiprt_camera_rays(const float *cam, const hiprt_ray_buffers *out, void *stream) {
    if (!cam || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (!out->d_ox || !out->d_oy || !out->d_oz || !out->d_tmin || !out->d_dx || !out->d_dy || !out->d_dz ||
        !out->d_tmax)
        return fail(HIPRT_ERR_ARG, "null ray plane");
    if (cam[10] < 1 || cam[11] < 1) return fail(HIPRT_ERR_ARG, "image size must be positive");
    rt::cam_frame f = frame_from(cam);
    rt::dev::ray_out o{out->d_ox, out->d_oy, out->d_oz, out->d_tmin, out->d_dx, out->d_dy, out->d_dz, out->d_tmax};
    HIP_TRY(rt::dev::launch_camera_rays(f, o, nullptr, nullptr, nullptr, false, static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_incoherent_rays(const float *sphere1, const float *sphere2, float r1, float r2, float r_max, uint64_t seed,
                          const hiprt_ray_buffers *out, uint32_t n, void *stream) {
    if (!sphere1 || !sphere2 || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (!(r1 >= 0) || !(r2 >= 0) || !(r_max >= 0) || !std::isfinite(r1) || !std::isfinite(r2) || !std::isfinite(r_max))
        return fail(HIPRT_ERR_ARG, "incoherence radii must be finite and non-negative");
    if (n == 0) return ok();
    if (!out->d_ox || !out->d_oy || !out->d_oz || !out->d_tmin || !out->d_dx || !out->d_dy || !out->d_dz ||
        !out->d_tmax)
        return fail(HIPRT_ERR_ARG, "null ray plane");
    rt::dev::ray_out o{out->d_ox, out->d_oy, out->d_oz, out->d_tmin, out->d_dx, out->d_dy, out->d_dz, out->d_tmax};
    HIP_TRY(rt::dev::launch_incoherent_rays(rt::vec3f{sphere1[0], sphere1[1], sphere1[2]},
                                            rt::vec3f{sphere2[0], sphere2[1], sphere2[2]}, r1, r2, r_max, seed, o, n,
                                            static_cast<hipStream_t>(stream)));
    return ok();
}

size_t hiprt_sort_keys_workspace(uint32_t n) {
    // keys_tmp, vals_tmp (n words each), the tile counts, the digit totals
    // and the item count, each 256-byte aligned
    const size_t words[4] = {n, n, 256ull * rt::dev::radix_tiles(n), 256ull + 64ull};
    size_t bytes = 0;
    for (size_t w : words) bytes += (4 * w + 255) & ~(size_t)255;
    return bytes;
}

int hiprt_sort_keys(uint32_t *d_keys, uint32_t *d_perm, uint32_t n, int bits, void *d_workspace, void *stream) {
    if (bits < 1 || bits > 32) return fail(HIPRT_ERR_ARG, "bits must be 1..32");
    if (n == 0) return ok();
    if (!d_keys || !d_perm || !d_workspace) return fail(HIPRT_ERR_ARG, "null argument");
    char *w = static_cast<char *>(d_workspace);
    auto take = [&w](size_t words) {
        uint32_t *p = reinterpret_cast<uint32_t *>(w);
        w += (4 * words + 255) & ~(size_t)255;
        return p;
    };
    rt::dev::radix_sort_buffers rb{};
    rb.keys_in = d_keys;
    rb.vals_out = d_perm;
    rb.keys_tmp = take(n);
    rb.vals_tmp = take(n);
    rb.counts = take(256ull * rt::dev::radix_tiles(n));
    rb.totals = take(256ull + 64ull);
    uint32_t *count = rb.totals + 256;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(rt::dev::launch_fill_u32(count, n, 1u, s));
    HIP_TRY(rt::dev::launch_radix_sort(rb, n, count, bits, s));
    return ok();
}

int hiprt_bounce_cosine(hiprt_scene sc, const hiprt_rays *rays, const hiprt_hits *hits, uint32_t n, uint64_t seed,
                        int32_t level, const hiprt_ray_buffers *out, void *stream) {
    if (!sc || !rays || !hits || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    if (n == 0) return ok();
    if (!rays->d_dx || !rays->d_dy || !rays->d_dz || !rays->d_tmax || !hits->d_t || !hits->d_u || !hits->d_v ||
        !hits->d_id)
        return fail(HIPRT_ERR_ARG, "null input plane");
    if (!out->d_ox || !out->d_oy || !out->d_oz || !out->d_tmin || !out->d_dx || !out->d_dy || !out->d_dz ||
        !out->d_tmax)
        return fail(HIPRT_ERR_ARG, "null output plane");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::ray_soa in{rays->d_ox, rays->d_oy, rays->d_oz, rays->d_tmin, rays->d_dx, rays->d_dy, rays->d_dz, rays->d_tmax};
    rt::hit_soa h{hits->d_t, hits->d_u, hits->d_v, hits->d_id};
    rt::dev::ray_out o{out->d_ox, out->d_oy, out->d_oz, out->d_tmin, out->d_dx, out->d_dy, out->d_dz, out->d_tmax};
    rt::dev::pt_scene ps = pt_scene_of(sc);
    HIP_TRY(rt::dev::launch_cosine_bounce(ps, in, h, o, seed, level, n, static_cast<hipStream_t>(stream)));
    return ok();
}

namespace {
bool full_rays(const hiprt_rays *r) {
    return r && r->d_ox && r->d_oy && r->d_oz && r->d_tmin && r->d_dx && r->d_dy && r->d_dz && r->d_tmax;
}
bool full_hits(const hiprt_hits *h) { return h && h->d_t && h->d_u && h->d_v && h->d_id; }
rt::ray_soa soa(const hiprt_rays *r) {
    return rt::ray_soa{r->d_ox, r->d_oy, r->d_oz, r->d_tmin, r->d_dx, r->d_dy, r->d_dz, r->d_tmax};
}
rt::hit_soa hsoa(const hiprt_hits *h) { return rt::hit_soa{h->d_t, h->d_u, h->d_v, h->d_id}; }
} // namespace

int hiprt_add_hitpoint_albedo(hiprt_scene sc, const hiprt_hits *hits, uint32_t n, float *d_framebuffer, void *stream) {
    if (!sc || !full_hits(hits) || (n && !d_framebuffer)) return fail(HIPRT_ERR_ARG, "null argument");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::pt_scene ps = pt_scene_of(sc);
    HIP_TRY(rt::dev::launch_add_albedo(ps, hsoa(hits), reinterpret_cast<float4 *>(d_framebuffer), n,
                                       static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_shadow_rays(hiprt_scene sc, const hiprt_rays *rays, const hiprt_hits *hits, uint32_t n, const float *light,
                      const hiprt_ray_buffers *out, void *stream) {
    if (!sc || !full_rays(rays) || !full_hits(hits) || !light || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (!out->d_ox || !out->d_oy || !out->d_oz || !out->d_tmin || !out->d_dx || !out->d_dy || !out->d_dz ||
        !out->d_tmax)
        return fail(HIPRT_ERR_ARG, "null output plane");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::pt_scene ps = pt_scene_of(sc);
    rt::dev::ray_out o{out->d_ox, out->d_oy, out->d_oz, out->d_tmin, out->d_dx, out->d_dy, out->d_dz, out->d_tmax};
    HIP_TRY(rt::dev::launch_shadow_rays(ps, soa(rays), hsoa(hits), rt::vec3f{light[0], light[1], light[2]}, o, n,
                                        static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_integrate_light(hiprt_scene sc, const hiprt_rays *cam_rays, const hiprt_hits *hits,
                          const hiprt_rays *shadow_rays, const uint8_t *d_occluded, uint32_t n, const float *light,
                          const float *col, float *d_framebuffer, void *stream) {
    if (!sc || !full_rays(cam_rays) || !full_hits(hits) || !full_rays(shadow_rays) || !light || !col ||
        (n && (!d_occluded || !d_framebuffer)))
        return fail(HIPRT_ERR_ARG, "null argument");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::pt_scene ps = pt_scene_of(sc);
    HIP_TRY(rt::dev::launch_integrate_light(ps, soa(cam_rays), hsoa(hits), soa(shadow_rays), d_occluded,
                                            rt::vec3f{light[0], light[1], light[2]},
                                            rt::vec3f{col[0], col[1], col[2]},
                                            reinterpret_cast<float4 *>(d_framebuffer), n,
                                            static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_scene_set_lights(hiprt_scene sc, const float *point_lights, uint32_t n_points, uint32_t *n_lights) {
    if (!sc || (n_points && !point_lights)) return fail(HIPRT_ERR_ARG, "null argument");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    // the emissive triangles, in triangle order, from the committed shading data
    std::vector<float> geo(12ull * sc->n_tris), mats(sc->mats.bytes / 4);
    if (sc->n_tris) HIP_TRY(hipMemcpy(geo.data(), sc->tri_geo.p, geo.size() * 4, hipMemcpyDeviceToHost));
    if (!mats.empty()) HIP_TRY(hipMemcpy(mats.data(), sc->mats.p, mats.size() * 4, hipMemcpyDeviceToHost));
    std::vector<rt::light_rec> recs;
    std::vector<float> f;
    auto bits = [](uint32_t u) {
        float x;
        std::memcpy(&x, &u, 4);
        return x;
    };
    for (uint32_t t = 0; t < sc->n_tris; ++t) {
        const float *g = &geo[12ull * t];
        uint32_t m;
        std::memcpy(&m, &g[3], 4);
        if (8ull * m + 8 > mats.size()) return fail(HIPRT_ERR_ARG, "triangle material id out of range");
        rt::vec3f e{mats[8ull * m + 4], mats[8ull * m + 5], mats[8ull * m + 6]};
        if (e == rt::v3(0.0f)) continue;
        rt::vec3f A{g[0], g[1], g[2]}, B{g[4], g[5], g[6]}, C{g[8], g[9], g[10]};
        recs.push_back(rt::light_rec{{A.x, A.y, A.z, bits(rt::kLightTriangle), B.x, B.y, B.z, rt::tri_area(A, B, C),
                                      C.x, C.y, C.z, bits(t), e.x, e.y, e.z, 0.0f}});
        f.push_back(rt::triangle_light_power(A, B, C, e));
    }
    for (uint32_t k = 0; k < n_points; ++k) {
        const float *p = point_lights + 6ull * k;
        rt::vec3f col{p[3], p[4], p[5]};
        recs.push_back(rt::light_rec{{p[0], p[1], p[2], bits(rt::kLightPoint), 0, 0, 0, 0, 0, 0, 0, 0, col.x, col.y,
                                      col.z, 0.0f}});
        f.push_back(rt::point_light_power(col));
    }
    // distribution_1d::build_cdf (libgi/discrete_distributions.cpp:25-38)
    const uint32_t n = (uint32_t)recs.size();
    std::vector<float> cdf(n + 1, 0.0f);
    for (uint32_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1];
    const float integral = cdf[n];
    for (uint32_t i = 1; i < n + 1; ++i) cdf[i] = integral > 0 ? cdf[i] / integral : (float)i / (float)n;
    sc->light_rec.release();
    sc->light_f.release();
    sc->light_cdf.release();
    sc->n_lights = 0;
    if (n) {
        HIP_TRY(sc->light_rec.upload(recs.data(), recs.size() * sizeof(rt::light_rec)));
        HIP_TRY(sc->light_f.upload(f.data(), f.size() * 4));
        HIP_TRY(sc->light_cdf.upload(cdf.data(), cdf.size() * 4));
    }
    sc->n_lights = n;
    sc->light_integral = integral;
    if (n_lights) *n_lights = n;
    return ok();
}

int hiprt_sample_direct(hiprt_scene sc, int32_t mode, const hiprt_rays *cam_rays, const hiprt_hits *cam_hits,
                        uint32_t n, uint64_t seed, const hiprt_ray_buffers *out, float *d_pdf, float *d_light_col,
                        float *d_framebuffer, void *stream) {
    if (!sc || !full_rays(cam_rays) || !full_hits(cam_hits) || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (mode < HIPRT_DIRECT_UNIFORM || mode > HIPRT_DIRECT_LIGHT) return fail(HIPRT_ERR_ARG, "unknown sampling mode");
    if (!out->d_ox || !out->d_oy || !out->d_oz || !out->d_tmin || !out->d_dx || !out->d_dy || !out->d_dz ||
        !out->d_tmax)
        return fail(HIPRT_ERR_ARG, "null output plane");
    if (n && (!d_pdf || !d_framebuffer)) return fail(HIPRT_ERR_ARG, "null pdf or framebuffer");
    if (mode == HIPRT_DIRECT_LIGHT && n && !d_light_col) return fail(HIPRT_ERR_ARG, "light mode needs d_light_col");
    if (mode == HIPRT_DIRECT_LIGHT && sc->n_lights == 0)
        return fail(HIPRT_ERR_ARG, "light sampling needs lights (hiprt_scene_set_lights found none)");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::pt_scene ps = pt_scene_of(sc);
    rt::dev::light_table lt{sc->light_rec.as<const float4>(), sc->light_f.as<const float>(),
                            sc->light_cdf.as<const float>(), sc->n_lights, sc->light_integral};
    rt::dev::ray_out o{out->d_ox, out->d_oy, out->d_oz, out->d_tmin, out->d_dx, out->d_dy, out->d_dz, out->d_tmax};
    HIP_TRY(rt::dev::launch_sample_direct(ps, lt, mode, soa(cam_rays), hsoa(cam_hits), seed, o, d_pdf, d_light_col,
                                          reinterpret_cast<float4 *>(d_framebuffer), n,
                                          static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_integrate_dir_sample(hiprt_scene sc, const hiprt_rays *cam_rays, const hiprt_hits *cam_hits,
                               const hiprt_rays *shadow_rays, const hiprt_hits *shadow_hits, const float *d_pdf,
                               uint32_t n, float *d_framebuffer, void *stream) {
    if (!sc || !full_rays(cam_rays) || !full_hits(cam_hits) || !full_rays(shadow_rays) || !full_hits(shadow_hits))
        return fail(HIPRT_ERR_ARG, "null argument");
    if (n && (!d_pdf || !d_framebuffer)) return fail(HIPRT_ERR_ARG, "null pdf or framebuffer");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::pt_scene ps = pt_scene_of(sc);
    HIP_TRY(rt::dev::launch_integrate_dir(ps, soa(cam_rays), hsoa(cam_hits), soa(shadow_rays), hsoa(shadow_hits),
                                          d_pdf, reinterpret_cast<float4 *>(d_framebuffer), n,
                                          static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_integrate_light_sample(hiprt_scene sc, const hiprt_rays *cam_rays, const hiprt_hits *cam_hits,
                                 const hiprt_rays *shadow_rays, const uint8_t *d_occluded, const float *d_pdf,
                                 const float *d_light_col, uint32_t n, float *d_framebuffer, void *stream) {
    if (!sc || !full_rays(cam_rays) || !full_hits(cam_hits) || !full_rays(shadow_rays))
        return fail(HIPRT_ERR_ARG, "null argument");
    if (n && (!d_occluded || !d_pdf || !d_light_col || !d_framebuffer))
        return fail(HIPRT_ERR_ARG, "null occlusion, pdf, light colour or framebuffer");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::pt_scene ps = pt_scene_of(sc);
    HIP_TRY(rt::dev::launch_integrate_light_sample(ps, soa(cam_rays), hsoa(cam_hits), soa(shadow_rays), d_occluded,
                                                   d_pdf, d_light_col, reinterpret_cast<float4 *>(d_framebuffer), n,
                                                   static_cast<hipStream_t>(stream)));
    return ok();
}

int hiprt_adjoint_closest(hiprt_scene sc, const hiprt_rays *rays, const hiprt_hits *hits, uint32_t n,
                          const float *d_grad_t, const float *d_grad_u, const float *d_grad_v, float *d_grad_vertices,
                          float *d_grad_origin, float *d_grad_dir, float *d_grad_corners, void *stream) {
    if (!sc || !full_rays(rays) || !hits || !hits->d_t || !hits->d_id) return fail(HIPRT_ERR_ARG, "null argument");
    if (n && (!d_grad_t || !d_grad_u || !d_grad_v)) return fail(HIPRT_ERR_ARG, "null upstream gradient");
    if (!sc->tri_geo.p || !sc->tri_vidx.p)
        return fail(HIPRT_ERR_UNSUPPORTED, "scene has no vertex data (commit it from a host scene)");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::dev::adjoint_args a{soa(rays),
                            rt::hit_soa{hits->d_t, hits->d_u, hits->d_v, hits->d_id},
                            sc->tri_geo.as<const float4>(),
                            sc->tri_vidx.as<const uint32_t>(),
                            d_grad_t,
                            d_grad_u,
                            d_grad_v,
                            d_grad_vertices,
                            d_grad_origin,
                            d_grad_dir,
                            d_grad_corners,
                            sc->n_tris};
    hipStream_t s = static_cast<hipStream_t>(stream);
    void *scratch = nullptr;
    if (d_grad_vertices && n) {
        // the stream's sort scratch, grown once its earlier launches are done
        hiprt_stream_state *st = sc->ctx->state(stream);
        if (!st) return fail(HIPRT_ERR_HIP, "cannot allocate the stream's work counters");
        const size_t bytes = rt::dev::adjoint_sort_bytes(n);
        if (st->adjoint.bytes < bytes) {
            HIP_TRY(hipStreamSynchronize(s));
            hipError_t e = st->adjoint.alloc(bytes);
            if (e != hipSuccess) {
                st->adjoint.release();
                return fail(HIPRT_ERR_HIP, std::string("adjoint sort scratch: ") + hipGetErrorString(e));
            }
        }
        scratch = st->adjoint.p;
    }
    HIP_TRY(rt::dev::launch_adjoint_closest(a, n, scratch, s));
    return ok();
}

int hiprt_scene_counts(hiprt_scene sc, uint32_t *n_vert, uint32_t *n_tris) {
    if (!sc) return fail(HIPRT_ERR_ARG, "null scene");
    if (n_vert) *n_vert = sc->n_vert;
    if (n_tris) *n_tris = sc->n_tris;
    return ok();
}

int hiprt_pt_create(hiprt_scene sc, const float *cam, const hiprt_pt_params *params, hiprt_pt *out) {
    if (!sc || !cam || !params || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (!sc->tri_geo.p) return fail(HIPRT_ERR_UNSUPPORTED, "scene has no shading data (commit it from a host scene)");
    if (cam[10] < 1 || cam[11] < 1) return fail(HIPRT_ERR_ARG, "image size must be positive");
    if (params->max_depth < 0 || params->max_depth > 64) return fail(HIPRT_ERR_ARG, "max_depth must be 0..64");
    if (!(params->rr_survive > 0.0f && params->rr_survive <= 1.0f)) return fail(HIPRT_ERR_ARG, "rr_survive must be in (0, 1]");
    double npix = (double)cam[10] * (double)cam[11];
    if (npix > 0x40000000) return fail(HIPRT_ERR_ARG, "image too large");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    std::unique_ptr<hiprt_pt_s> p(new (std::nothrow) hiprt_pt_s);
    if (!p) return fail(HIPRT_ERR_ARG, "out of host memory");
    p->sc = sc;
    p->f = frame_from(cam);
    const uint32_t samples = params->samples > 1 ? (uint32_t)params->samples : 1u;
    if (samples > 1 && !params->batched) return fail(HIPRT_ERR_ARG, "samples in flight need the batched schedule");
    if (npix * samples > 0x40000000) return fail(HIPRT_ERR_ARG, "image too large for that many samples in flight");
    p->n_px = (uint32_t)npix;
    p->samples = samples;
    if (params->regen_pool < 0) return fail(HIPRT_ERR_ARG, "regen_pool must be >= 0");
    if (params->regen_pool > 0) {
        if (!params->batched) return fail(HIPRT_ERR_ARG, "path regeneration needs the batched schedule");
        if (samples >= (1u << 24)) return fail(HIPRT_ERR_ARG, "path regeneration takes fewer than 2^24 samples");
        p->lanes = std::min((uint32_t)params->regen_pool, samples);
    }
    // path slots: the regeneration pool, or every sample's paths at once
    p->n = (uint32_t)npix * (p->lanes ? p->lanes : samples);
    p->pp = rt::dev::pt_params{params->max_depth, params->rr_start, params->rr_survive, (uint32_t)npix,
                               params->emitters_reflect ? 1 : 0};
    // samples in flight: paths numbered pixel-major, so a pixel's samples
    // share wavefronts from the camera level on
    if (!p->lanes && samples > 1) p->pp.px_samples = samples;
    p->seed = params->seed;
    p->batched = params->batched != 0;
    p->jitter = params->pixel_centre == 0;
    int tracer = params->tracer;
    p->alpha = tracer >= 0 && (tracer & HIPRT_TRACER_ALPHA) != 0;
    if (tracer >= 0) tracer &= ~HIPRT_TRACER_ALPHA;
    p->tracer = tracer >= 0 ? tracer : (p->batched ? HIPRT_TRACER_DYNAMIC_WHILE_WHILE : HIPRT_TRACER_WHILE_WHILE);
    if (p->tracer >= HIPRT_TRACER_COUNT) return fail(HIPRT_ERR_ARG, "unknown tracer id");
    if (p->batched && p->tracer == HIPRT_TRACER_WHILE_WHILE)
        return fail(HIPRT_ERR_ARG, "the batched schedule needs a persistent tracer");
    if (!p->batched && p->tracer != HIPRT_TRACER_WHILE_WHILE)
        return fail(HIPRT_ERR_ARG, "the plain schedule uses the while-while tracer");
    if (!p->batched && sc->desc.bvh != rt::BVH2)
        return fail(HIPRT_ERR_UNSUPPORTED, "the plain schedule needs a binary BVH scene");
    const int stack = (params->stack_config > 0 && params->stack_config < rt::dev::kStackConfigs) ? params->stack_config : 0;
    if (!p->batched && sc->stack_need > rt::dev::kPlainStackCap - 1)
        return fail(HIPRT_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    uint32_t n = p->n;
    const bool regen = p->lanes > 0;
    p->sorted = p->batched && params->ray_order == 0 && p->pp.max_depth > 0;
    if (!regen) {
        HIP_TRY(p->rays[0].alloc(n));
        if (p->batched && !p->sorted) {   // the sorting form carries path ids and throughputs in its records
            HIP_TRY(p->paths[0].alloc(4ull * n));
            HIP_TRY(p->rays[1].alloc(n));
            HIP_TRY(p->paths[1].alloc(4ull * n));
        }
        if (!p->sorted) HIP_TRY(p->thr.alloc(12ull * n));
        if (samples > 1) HIP_TRY(p->rad_paths.alloc(12ull * n));
    } else {
        // records in both parities, queue lengths, per-slot radiance
        for (int k = 0; k < 2; ++k) HIP_TRY(p->rec[k].alloc(16ull * rt::dev::kRayRecord * n));
        HIP_TRY(p->qcount.alloc(8));
        if (p->lanes > 1) HIP_TRY(p->acc.alloc(12ull * n));
        HIP_TRY(hipHostMalloc((void **)&p->host_q, 8));
        for (hipEvent_t &e : p->regen_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    HIP_TRY(p->hits.alloc(16ull * n));
    float *hb = p->hits.as<float>();
    p->h = rt::hit_soa{hb, hb + n, hb + 2ull * n, reinterpret_cast<uint32_t *>(hb + 3ull * n)};
    HIP_TRY(p->rad.alloc(12ull * p->n_px));
    HIP_TRY(hipMemset(p->rad.p, 0, 12ull * p->n_px));
    HIP_TRY(p->counts.alloc(4ull * (p->pp.max_depth + 2)));
    HIP_TRY(hipMemset(p->counts.p, 0, 4ull * (p->pp.max_depth + 2)));
    HIP_TRY(p->totals.alloc(8ull * (p->pp.max_depth + 1)));
    HIP_TRY(hipMemset(p->totals.p, 0, 8ull * (p->pp.max_depth + 1)));
    HIP_TRY(p->frame_seed.alloc(8ull * samples));
    HIP_TRY(p->work.alloc(rt::dev::kCounterBytes));
    p->pcfg.speculative = p->tracer >= HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE;
    p->pcfg.refill = p->tracer == HIPRT_TRACER_DYNAMIC_WHILE_WHILE;
    p->pcfg.refill_thr = params->refill_threshold > 0 ? std::min(params->refill_threshold, 64) : 40;
    p->refill_coherent = params->refill_coherent > 0 ? std::min(params->refill_coherent, 64) : 16;
    p->pcfg.blocks = 0;
    p->pcfg.stack = stack;
    p->pcfg.chunks = params->work_chunks == 1 ? 1 : 8;
    p->pcfg.park = (params->park_leaves > 0 && params->park_leaves < rt::dev::kParkClosest) ? params->park_leaves
                                                                                          : rt::dev::kParkClosest;
    if (p->sorted) {
        if (!regen)
            for (int k = 0; k < 2; ++k) HIP_TRY(p->rec[k].alloc(16ull * rt::dev::kRayRecord * n));
        HIP_TRY(p->keys.alloc(4ull * n));
        HIP_TRY(p->keys_tmp.alloc(4ull * n));
        HIP_TRY(p->vals_tmp.alloc(4ull * n));
        HIP_TRY(p->perm.alloc(4ull * n));
        HIP_TRY(p->sort_counts.alloc(4ull * 256 * rt::dev::radix_tiles(n)));
        HIP_TRY(p->sort_totals.alloc(4ull * 256));
        p->grid = grid_of(sc->host);
    }
    if (p->batched) {
        // Resolve the persistent grid size outside the capture, then record
        // the frame once.  The graph never clears the accumulator: frame 0
        // after a reset is rendered eagerly with clear_radiance set.
        if (params->blocks > 0) p->pcfg.blocks = params->blocks;
        else HIP_TRY(rt::dev::persistent_grid(sc->desc, false, p->pcfg, &p->pcfg.blocks));
        // the graph keeps the spill buffer's address: allocate it for the grid now
        p->pcfg.spill_cap = rt::dev::persistent_spill_entries(sc->desc.bvh, stack, sc->stack_need);
        if (p->pcfg.spill_cap) {
            HIP_TRY(p->spill.alloc(rt::dev::persistent_spill_bytes(p->pcfg.blocks, p->pcfg.spill_cap)));
            p->pcfg.spill = p->spill.as<int32_t>();
            p->pcfg.spill_bytes = p->spill.bytes;
        }
        HIP_TRY(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
        hipError_t rec = regen ? p->record_regen_chunk(p->cap) : p->record(p->cap, false);
        hipGraph_t g = nullptr;
        hipError_t end = hipStreamEndCapture(p->cap, &g);
        if (rec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(HIPRT_ERR_HIP, std::string("recording the frame graph: ") + hipGetErrorString(rec));
        }
        HIP_TRY(end);
        hipError_t inst = hipGraphInstantiate(regen ? &p->regen_exec : &p->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(inst);
    }
    *out = p.release();
    return ok();
}

int hiprt_pt_render(hiprt_pt pt, uint32_t frame, void *stream) {
    if (!pt) return fail(HIPRT_ERR_ARG, "null path tracer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipSetDevice(pt->sc->ctx->device));
    // render call `frame` traces the samples of frames frame*samples .. +samples-1
    HIP_TRY(rt::dev::launch_set_seeds(pt->frame_seed.as<uint64_t>(),
                                      pt->seed * 0x100000001B3ULL + (uint64_t)frame * pt->samples, pt->samples, s));
    if (pt->lanes) return pt->render_regen(s);
    if (pt->batched && !pt->clear_next) {
        HIP_TRY(hipGraphLaunch(pt->exec, s));
    } else {
        HIP_TRY(pt->record(s, pt->clear_next));
    }
    pt->clear_next = false;
    return ok();
}

int hiprt_pt_autotune(hiprt_scene bvh2_scene, hiprt_scene bvh4_scene, const float *cam,
                      const hiprt_pt_params *base, int32_t max_samples, int32_t trial_frames,
                      hiprt_pt_tune_result *out) {
    if (!bvh2_scene || !cam || !base || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (bvh2_scene->desc.bvh != rt::BVH2) return fail(HIPRT_ERR_ARG, "bvh2_scene is not a binary-tree scene");
    if (bvh4_scene && bvh4_scene->desc.bvh != rt::BVH4) return fail(HIPRT_ERR_ARG, "bvh4_scene is not a 4-wide scene");
    const int frames = trial_frames > 0 ? trial_frames : 2;
    const int max_s = max_samples > 1 ? std::min(max_samples, 64) : 1;
    HIP_TRY(hipSetDevice(bvh2_scene->ctx->device));
    struct events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    HIP_TRY(hipEventCreate(&ev.e[0]));
    HIP_TRY(hipEventCreate(&ev.e[1]));
    int count = 0;
    // rays per millisecond of one candidate (negative on failure)
    auto trial = [&](int layout, int tracer, int stack, int samples, int park, float *rate) -> int {
        hiprt_pt_params p = *base;
        p.batched = 1;
        p.tracer = tracer | (base->tracer >= 0 ? (base->tracer & HIPRT_TRACER_ALPHA) : 0);
        p.stack_config = stack;
        p.samples = samples;
        p.park_leaves = park;
        p.regen_pool = 0;
        hiprt_pt pt = nullptr;
        int st = hiprt_pt_create(layout == HIPRT_BVH4 ? bvh4_scene : bvh2_scene, cam, &p, &pt);
        if (st != HIPRT_OK) return st;
        struct destroy_at_exit {
            hiprt_pt p;
            ~destroy_at_exit() { (void)hiprt_pt_destroy(p); }
        } own{pt};
        const int levels = p.max_depth + 1;
        std::vector<uint64_t> t0(levels), t1(levels);
        if ((st = hiprt_pt_render(pt, 0, nullptr)) != HIPRT_OK) return st;
        if ((st = hiprt_pt_ray_totals(pt, t0.data(), levels)) != HIPRT_OK) return st;
        HIP_TRY(hipEventRecord(ev.e[0], nullptr));
        for (int f = 1; f <= frames; ++f)
            if ((st = hiprt_pt_render(pt, (uint32_t)f, nullptr)) != HIPRT_OK) return st;
        HIP_TRY(hipEventRecord(ev.e[1], nullptr));
        HIP_TRY(hipEventSynchronize(ev.e[1]));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
        if ((st = hiprt_pt_ray_totals(pt, t1.data(), levels)) != HIPRT_OK) return st;
        double rays = 0;
        for (int L = 0; L < levels; ++L) rays += (double)(t1[L] - t0[L]);
        ++count;
        *rate = ms > 0 ? (float)(rays / ms) : 0.0f;
        return HIPRT_OK;
    };
    hiprt_pt_tune_result best{};
    best.mrays_per_s = -1;
    auto consider = [&](int layout, int tracer, int stack, int samples, int park) -> int {
        float rate = 0;
        int st = trial(layout, tracer, stack, samples, park, &rate);
        if (st != HIPRT_OK) return st;
        if (rate / 1e3f > best.mrays_per_s)
            best = hiprt_pt_tune_result{tracer, stack, samples, layout, park, rate / 1e3f, 0};
        return HIPRT_OK;
    };
    // 1. tracer x layout x stack configuration, one sample
    for (int tracer : {HIPRT_TRACER_DYNAMIC_WHILE_WHILE, HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE}) {
        for (int stack : {0, 3, 4})
            if (int st = consider(HIPRT_BVH2, tracer, stack, 1, rt::dev::kParkClosest)) return st;
        if (bvh4_scene)
            for (int stack : {0, 3})
                if (int st = consider(HIPRT_BVH4, tracer, stack, 1, rt::dev::kParkClosest)) return st;
    }
    // 2. samples in flight for the winner and for the binary tree with stack 4
    const hiprt_pt_tune_result one = best;
    for (int s = 2; s <= max_s; s *= 2) {
        if (int st = consider(one.bvh_format, one.tracer, one.stack_config, s, rt::dev::kParkClosest)) return st;
        if (!(one.bvh_format == HIPRT_BVH2 && one.stack_config == 4))
            if (int st = consider(HIPRT_BVH2, one.tracer, 4, s, rt::dev::kParkClosest)) return st;
    }
    // 3. parked leaves
    const hiprt_pt_tune_result win = best;
    for (int park = 1; park < rt::dev::kParkClosest; ++park)
        if (int st = consider(win.bvh_format, win.tracer, win.stack_config, win.samples, park)) return st;
    best.candidates = count;
    *out = best;
    return ok();
}

int hiprt_pt_ray_counts(hiprt_pt pt, uint32_t *counts, int32_t n_levels) {
    if (!pt || !counts) return fail(HIPRT_ERR_ARG, "null argument");
    int levels = pt->pp.max_depth + 1;
    if (n_levels < levels) return fail(HIPRT_ERR_ARG, "counts array shorter than max_depth + 1");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(counts, pt->counts.p, 4ull * levels, hipMemcpyDeviceToHost));
    return ok();
}

int hiprt_pt_ray_totals(hiprt_pt pt, uint64_t *totals, int32_t n_levels) {
    if (!pt || !totals) return fail(HIPRT_ERR_ARG, "null argument");
    int levels = pt->pp.max_depth + 1;
    if (n_levels < levels) return fail(HIPRT_ERR_ARG, "totals array shorter than max_depth + 1");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(totals, pt->totals.p, 8ull * levels, hipMemcpyDeviceToHost));
    return ok();
}

int hiprt_pt_profile(hiprt_pt pt, uint32_t frame, void *stream, float *ms_per_level, uint64_t *fetch_stats,
                     uint32_t *rays_per_level, int32_t n_levels) {
    if (!pt) return fail(HIPRT_ERR_ARG, "null path tracer");
    int levels = pt->pp.max_depth + 1;
    if (n_levels < levels) return fail(HIPRT_ERR_ARG, "arrays shorter than max_depth + 1");
    if (pt->lanes) return fail(HIPRT_ERR_UNSUPPORTED, "path regeneration mixes levels in every launch");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipSetDevice(pt->sc->ctx->device));
    std::vector<hipEvent_t> ev(2 * levels, nullptr);
    struct ev_guard {
        std::vector<hipEvent_t> &v;
        ~ev_guard() {
            for (hipEvent_t e : v)
                if (e) (void)hipEventDestroy(e);
        }
    } guard{ev};
    for (hipEvent_t &e : ev) HIP_TRY(hipEventCreate(&e));
    uint64_t seed = pt->seed * 0x100000001B3ULL + (uint64_t)frame * pt->samples;
    HIP_TRY(rt::dev::launch_set_seeds(pt->frame_seed.as<uint64_t>(), seed, pt->samples, s));
    HIP_TRY(pt->record(s, pt->clear_next, ev.data(), nullptr));
    pt->clear_next = false;
    HIP_TRY(hipStreamSynchronize(s));
    if (ms_per_level)
        for (int L = 0; L < levels; ++L) HIP_TRY(hipEventElapsedTime(&ms_per_level[L], ev[2 * L], ev[2 * L + 1]));
    if (rays_per_level) HIP_TRY(hipMemcpy(rays_per_level, pt->counts.p, 4ull * levels, hipMemcpyDeviceToHost));
    if (fetch_stats) {
        dbuf st;
        HIP_TRY(st.alloc(16ull * levels));
        HIP_TRY(hipMemset(st.p, 0, 16ull * levels));
        HIP_TRY(rt::dev::launch_set_seeds(pt->frame_seed.as<uint64_t>(), seed, pt->samples, s));
        HIP_TRY(pt->record(s, false, nullptr, st.as<unsigned long long>()));
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(fetch_stats, st.p, 16ull * levels, hipMemcpyDeviceToHost));
    }
    return ok();
}

int hiprt_pt_memory(hiprt_pt pt, uint64_t *bytes) {
    if (!pt || !bytes) return fail(HIPRT_ERR_ARG, "null argument");
    const dbuf *bufs[] = {&pt->rays[0].mem, &pt->rays[1].mem, &pt->hits,     &pt->paths[0],  &pt->paths[1],
                          &pt->thr,         &pt->rad,         &pt->counts,   &pt->totals,    &pt->frame_seed,
                          &pt->work,        &pt->spill,       &pt->rad_paths, &pt->rec[0],   &pt->rec[1],
                          &pt->keys,        &pt->keys_tmp,    &pt->vals_tmp, &pt->perm,      &pt->sort_counts,
                          &pt->sort_totals, &pt->qcount,      &pt->acc};
    uint64_t sum = 0;
    for (const dbuf *b : bufs) sum += b->bytes;
    *bytes = sum;
    return ok();
}

int hiprt_pt_image(hiprt_pt pt, float *rgb) {
    if (!pt || !rgb) return fail(HIPRT_ERR_ARG, "null argument");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(rgb, pt->rad.p, 12ull * pt->n_px, hipMemcpyDeviceToHost));
    return ok();
}

int hiprt_pt_reset(hiprt_pt pt, void *stream) {
    if (!pt) return fail(HIPRT_ERR_ARG, "null path tracer");
    (void)stream;
    pt->clear_next = true;
    return ok();
}

int hiprt_pt_destroy(hiprt_pt pt) {
    if (!pt) return fail(HIPRT_ERR_ARG, "null path tracer");
    (void)hipSetDevice(pt->sc->ctx->device);
    (void)hipDeviceSynchronize();
    delete pt;
    return ok();
}

} // extern "C"
