// Implementation of the C ABI declared in include/illum_hiprt.h.
//
// Every entry point validates its handles and sizes on the host before any
// kernel sees them: a BVH is refused unless its child indices strictly
// increase along every edge (so traversal terminates) and its depth fits the
// chosen kernel's stack, and ray/hit pointers must be non-null for every
// plane the kernel reads or writes.
#include "capi_internal.h"

#include "device/refit_kernels.h"

#include "core/raygen.h"
#include "device/kernels.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

namespace {
thread_local std::string g_err;

// Scratch of the allocation-free trace (hiprt_trace_closest_noalloc): the
// persistent kernels' work counters followed by a traversal-stack spill
// column.  One copy per device, held for the life of the process.  Host code
// never uses the array's name as an address: hiprt_create resolves each
// device's copy once with hipGetSymbolAddress and keeps it in
// g_scratch_addr, and only that address reaches kernels.
constexpr size_t kNoAllocSpillBytes = size_t(16) << 20;
constexpr size_t kNoAllocScratchBytes = rt::dev::kCounterBytes + kNoAllocSpillBytes;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void *g_scratch_addr[kMaxDevices] = {};
} // namespace

static __device__ __align__(256) uint32_t scratch[kNoAllocScratchBytes / 4];

namespace hiprt_detail {
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
rt::scene_desc desc_for(const hiprt_scene_s *sc, bool alpha) {
    rt::scene_desc d = sc->desc;
    d.alpha = (alpha && d.alpha_tri) ? 1 : 0;
    return d;
}

hipError_t attach_spill(hiprt_stream_state *st, hipStream_t s, const rt::scene_desc &d, bool any, bool multi,
                        int need, rt::dev::persistent_cfg &cfg, int spill_limit) {
    cfg.spill = nullptr;
    cfg.spill_cap = rt::dev::persistent_spill_entries(d.bvh, cfg.stack, need);
    if (spill_limit > 0 && cfg.spill_cap > (uint32_t)spill_limit) cfg.spill_cap = (uint32_t)spill_limit;
    cfg.spill_bytes = 0;
    if (cfg.spill_cap == 0) return hipSuccess;
    int blocks = 0;
    hipError_t e = rt::dev::persistent_grid(d, any, cfg, &blocks, multi);
    if (e != hipSuccess) return e;
    size_t bytes = rt::dev::persistent_spill_bytes(blocks, cfg.spill_cap);
    if (st->spill.bytes < bytes) {
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        e = st->spill.alloc(bytes);
        if (e != hipSuccess) {
            st->spill.release();
            return e;
        }
    }
    cfg.spill = st->spill.as<int32_t>();
    cfg.spill_bytes = st->spill.bytes;
    return hipSuccess;
}

} // namespace hiprt_detail

using hiprt_detail::dbuf;
using hiprt_detail::fail;
using hiprt_detail::ok;

hiprt_stream_state *hiprt_context_s::state(void *stream) {
    std::lock_guard<std::mutex> g(lock);
    auto it = streams.find(stream);
    if (it != streams.end()) return it->second.get();
    std::unique_ptr<hiprt_stream_state> st(new (std::nothrow) hiprt_stream_state);
    if (!st) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    if (st->counter.alloc(rt::dev::kCounterBytes) != hipSuccess) return nullptr;
    if (hipMemset(st->counter.p, 0, rt::dev::kCounterBytes) != hipSuccess) return nullptr;
    hiprt_stream_state *p = st.get();
    streams.emplace(stream, std::move(st));
    return p;
}

extern "C" {

const char *hiprt_last_error(void) { return g_err.c_str(); }
const char *hiprt_version(void) { return "illum-hiprt 0.1 (gfx950)"; }

// ------------------------------------------------------------------- host rays

int hiprt_camera_rays_host(const float *cam, float *const *planes) {
    if (!cam || !planes) return fail(HIPRT_ERR_ARG, "null argument");
    for (int k = 0; k < 8; ++k)
        if (!planes[k]) return fail(HIPRT_ERR_ARG, "null ray plane");
    int w = (int)cam[10], h = (int)cam[11];
    if (w <= 0 || h <= 0) return fail(HIPRT_ERR_ARG, "image size must be positive");
    rt::cam_frame f = rt::make_cam_frame(rt::vec3f{cam[0], cam[1], cam[2]}, rt::vec3f{cam[3], cam[4], cam[5]},
                                         rt::vec3f{cam[6], cam[7], cam[8]}, cam[9], w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            rt::vec3f d = rt::cam_dir(f, x, y);
            planes[0][i] = f.pos.x;
            planes[1][i] = f.pos.y;
            planes[2][i] = f.pos.z;
            planes[3][i] = rt::kRayEps;
            planes[4][i] = d.x;
            planes[5][i] = d.y;
            planes[6][i] = d.z;
            planes[7][i] = FLT_MAX;
        }
    return ok();
}

// ------------------------------------------------------------------ host scene

int hiprt_scene_generate(int config, hiprt_host_scene *out, float *cam) {
    if (!out) return fail(HIPRT_ERR_ARG, "out is null");
    if (config < 1 || config > 3) return fail(HIPRT_ERR_ARG, "config must be 1, 2 or 3");
    auto *h = new (std::nothrow) hiprt_host_scene_s;
    if (!h) return fail(HIPRT_ERR_ARG, "out of host memory");
    rt::workload w = rt::make_workload(config);
    h->sc = std::move(w.sc);
    if (cam) {
        const rt::camera &c = w.cam;
        float v[15] = {c.pos.x, c.pos.y, c.pos.z, c.dir.x, c.dir.y, c.dir.z, c.up.x, c.up.y, c.up.z,
                       c.fovy, (float)c.w, (float)c.h, w.light.x, w.light.y, w.light.z};
        std::memcpy(cam, v, sizeof v);
    }
    *out = h;
    return ok();
}

int hiprt_scene_random(uint32_t n_tris, float sigma, uint64_t seed, hiprt_host_scene *out) {
    if (!out) return fail(HIPRT_ERR_ARG, "out is null");
    auto *h = new (std::nothrow) hiprt_host_scene_s;
    if (!h) return fail(HIPRT_ERR_ARG, "out of host memory");
    h->sc = rt::gen_random_tris(n_tris, sigma, seed);
    *out = h;
    return ok();
}

int hiprt_scene_from_arrays(const float *pos, uint32_t n_vert, const uint32_t *idx, uint32_t n_tris,
                            const uint32_t *mat, hiprt_host_scene *out) {
    if (!out || (n_vert && !pos) || (n_tris && !idx)) return fail(HIPRT_ERR_ARG, "null array");
    for (uint64_t k = 0; k < 3ull * n_tris; ++k)
        if (idx[k] >= n_vert) return fail(HIPRT_ERR_ARG, "triangle index out of range");
    auto *h = new (std::nothrow) hiprt_host_scene_s;
    if (!h) return fail(HIPRT_ERR_ARG, "out of host memory");
    h->sc.pos.resize(n_vert);
    for (uint32_t i = 0; i < n_vert; ++i) h->sc.pos[i] = rt::vec3f{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    h->sc.idx.assign(idx, idx + 3ull * n_tris);
    uint32_t max_mat = 0;
    if (mat) {
        h->sc.mat.assign(mat, mat + n_tris);
        for (uint32_t m : h->sc.mat) max_mat = m > max_mat ? m : max_mat;
    } else {
        h->sc.mat.assign(n_tris, 0);
    }
    h->sc.mats.assign(max_mat + 1, rt::material{rt::v3(0.8f), rt::v3(0.0f)});
    *out = h;
    return ok();
}

int hiprt_scene_load_obj(const char *path, hiprt_host_scene *out) {
    if (!path || !out) return fail(HIPRT_ERR_ARG, "null argument");
    auto *h = new (std::nothrow) hiprt_host_scene_s;
    if (!h) return fail(HIPRT_ERR_ARG, "out of host memory");
    std::string err;
    if (!rt::read_obj(path, h->sc, err)) {
        delete h;
        return fail(HIPRT_ERR_IO, err);
    }
    *out = h;
    return ok();
}

int hiprt_scene_add_obj(hiprt_host_scene *scene, const char *path, const char *name, const float *trafo) {
    if (!scene || !path) return fail(HIPRT_ERR_ARG, "null argument");
    rt::scene loaded;
    std::string err;
    if (!rt::read_obj(path, loaded, err, trafo)) return fail(HIPRT_ERR_IO, err);
    const std::string prefix = name ? name : "";
    if (*scene) {
        rt::append_scene((*scene)->sc, loaded, prefix);
        return ok();
    }
    auto *h = new (std::nothrow) hiprt_host_scene_s;
    if (!h) return fail(HIPRT_ERR_ARG, "out of host memory");
    rt::append_scene(h->sc, loaded, prefix);
    *scene = h;
    return ok();
}

int hiprt_scene_write_obj(hiprt_host_scene s, const char *path) {
    if (!s || !path) return fail(HIPRT_ERR_ARG, "null argument");
    if (!rt::write_obj(s->sc, path)) return fail(HIPRT_ERR_IO, std::string("cannot write ") + path);
    return ok();
}

int hiprt_scene_info(hiprt_host_scene s, uint32_t *n_vert, uint32_t *n_tris) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    if (n_vert) *n_vert = (uint32_t)s->sc.pos.size();
    if (n_tris) *n_tris = s->sc.n_tris();
    return ok();
}

int hiprt_scene_materials(hiprt_host_scene s, float *table, uint32_t *n_materials) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    if (n_materials) *n_materials = (uint32_t)s->sc.mats.size();
    if (table)
        for (size_t m = 0; m < s->sc.mats.size(); ++m) {
            const rt::material &mt = s->sc.mats[m];
            float v[8] = {mt.albedo.x, mt.albedo.y, mt.albedo.z, 0.0f, mt.emissive.x, mt.emissive.y, mt.emissive.z, 0.0f};
            std::memcpy(table + 8 * m, v, sizeof v);
        }
    return ok();
}

int hiprt_scene_material_name(hiprt_host_scene s, uint32_t material, char *name, uint32_t size) {
    if (!s || !name || size == 0) return fail(HIPRT_ERR_ARG, "null argument");
    if (material >= s->sc.mats.size()) return fail(HIPRT_ERR_ARG, "material id out of range");
    const std::string n = material < s->sc.mat_names.size() ? s->sc.mat_names[material] : std::string();
    if (n.size() + 1 > size) return fail(HIPRT_ERR_ARG, "name buffer too small");
    std::memcpy(name, n.c_str(), n.size() + 1);
    return ok();
}

int hiprt_scene_set_material(hiprt_host_scene s, uint32_t material, const float *albedo, const float *emissive) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    if (material >= s->sc.mats.size()) return fail(HIPRT_ERR_ARG, "material id out of range");
    rt::material &m = s->sc.mats[material];
    if (albedo) m.albedo = rt::vec3f{albedo[0], albedo[1], albedo[2]};
    if (emissive) m.emissive = rt::vec3f{emissive[0], emissive[1], emissive[2]};
    return ok();
}

int hiprt_scene_set_texcoords(hiprt_host_scene s, const float *tc, uint32_t n_vert) {
    if (!s || (n_vert && !tc)) return fail(HIPRT_ERR_ARG, "null argument");
    if (n_vert != s->sc.pos.size()) return fail(HIPRT_ERR_ARG, "texture coordinate count differs from the vertex count");
    s->sc.tc.assign(tc, tc + 2ull * n_vert);
    return ok();
}

int hiprt_scene_get_texcoords(hiprt_host_scene s, float *tc, uint32_t *n_vert) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    const size_t nv = s->sc.pos.size();
    if (n_vert) *n_vert = (uint32_t)nv;
    if (tc) {
        for (size_t i = 0; i < 2 * nv; ++i) tc[i] = i < s->sc.tc.size() ? s->sc.tc[i] : 0.0f;
    }
    return ok();
}

int hiprt_scene_set_positions(hiprt_host_scene s, const float *pos, uint32_t n_vert) {
    if (!s || (n_vert && !pos)) return fail(HIPRT_ERR_ARG, "null argument");
    if (n_vert != s->sc.pos.size()) return fail(HIPRT_ERR_ARG, "position count differs from the vertex count");
    for (uint32_t i = 0; i < n_vert; ++i) s->sc.pos[i] = rt::vec3f{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    return ok();
}

int hiprt_scene_set_normals(hiprt_host_scene s, const float *normals, uint32_t n_vert) {
    if (!s || (n_vert && !normals)) return fail(HIPRT_ERR_ARG, "null argument");
    if (n_vert && n_vert != s->sc.pos.size()) return fail(HIPRT_ERR_ARG, "normal count differs from the vertex count");
    s->sc.norm.assign(normals, normals + 3ull * n_vert);
    return ok();
}

int hiprt_scene_get_normals(hiprt_host_scene s, float *normals, uint32_t *n_vert) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    const bool has = !s->sc.norm.empty() && s->sc.norm.size() == 3 * s->sc.pos.size();
    if (n_vert) *n_vert = has ? (uint32_t)s->sc.pos.size() : 0u;
    if (normals && has) std::memcpy(normals, s->sc.norm.data(), s->sc.norm.size() * 4);
    return ok();
}

int hiprt_scene_set_texture(hiprt_host_scene s, uint32_t material, const float *rgba, uint32_t w, uint32_t h) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    if (material >= s->sc.mats.size()) return fail(HIPRT_ERR_ARG, "material id out of range");
    rt::material &m = s->sc.mats[material];
    if (!rgba) {
        m.tex = -1;
        return ok();
    }
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28)) return fail(HIPRT_ERR_ARG, "texture size must be 1..2^28 texels");
    rt::texture t;
    t.w = w;
    t.h = h;
    t.rgba.assign(rgba, rgba + 4ull * w * h);
    if (m.tex >= 0) {
        s->sc.texs[(size_t)m.tex] = std::move(t);
    } else {
        m.tex = (int32_t)s->sc.texs.size();
        s->sc.texs.push_back(std::move(t));
    }
    return ok();
}

int hiprt_scene_copy(hiprt_host_scene s, float *pos, uint32_t *idx, uint32_t *mat) {
    if (!s) return fail(HIPRT_ERR_ARG, "null scene");
    if (pos)
        for (size_t i = 0; i < s->sc.pos.size(); ++i) {
            pos[3 * i] = s->sc.pos[i].x;
            pos[3 * i + 1] = s->sc.pos[i].y;
            pos[3 * i + 2] = s->sc.pos[i].z;
        }
    if (idx && !s->sc.idx.empty()) std::memcpy(idx, s->sc.idx.data(), s->sc.idx.size() * 4);
    if (mat && !s->sc.mat.empty()) std::memcpy(mat, s->sc.mat.data(), s->sc.mat.size() * 4);
    return ok();
}

void hiprt_scene_free(hiprt_host_scene s) { delete s; }

// -------------------------------------------------------------------- host BVH

int hiprt_bvh_build(hiprt_host_scene s, const hiprt_bvh_params *params, hiprt_bvh *out) {
    if (!s || !out) return fail(HIPRT_ERR_ARG, "null argument");
    rt::bvh_params p = rt::gpu_default_params();
    if (params) {
        if (params->split < 0 || params->split > 2) return fail(HIPRT_ERR_ARG, "split must be 0 (om), 1 (sm) or 2 (sah)");
        if (params->max_tris < 1 || params->planes < 1) return fail(HIPRT_ERR_ARG, "max_tris and planes must be >= 1");
        p.split = params->split == 0 ? rt::split_kind::om : (params->split == 1 ? rt::split_kind::sm : rt::split_kind::sah);
        p.max_tris = params->max_tris;
        p.planes = params->planes;
        p.esc = params->esc != 0;
    }
    auto *h = new (std::nothrow) hiprt_bvh_s;
    if (!h) return fail(HIPRT_ERR_ARG, "out of host memory");
    h->b = rt::build_bvh(s->sc, p);
    h->st = rt::compute_stats(h->b);
    *out = h;
    return ok();
}

int hiprt_bvh_info(hiprt_bvh b, uint32_t *n_nodes, uint32_t *n_index, uint32_t *root, uint32_t *depth) {
    if (!b) return fail(HIPRT_ERR_ARG, "null bvh");
    if (n_nodes) *n_nodes = (uint32_t)b->b.nodes.size();
    if (n_index) *n_index = (uint32_t)b->b.index.size();
    if (root) *root = b->b.root;
    if (depth) *depth = b->st.depth;
    return ok();
}

int hiprt_bvh_copy(hiprt_bvh b, void *nodes, uint32_t *index) {
    if (!b) return fail(HIPRT_ERR_ARG, "null bvh");
    if (nodes && !b->b.nodes.empty()) std::memcpy(nodes, b->b.nodes.data(), b->b.nodes.size() * sizeof(rt::bbvh_node));
    if (index && !b->b.index.empty()) std::memcpy(index, b->b.index.data(), b->b.index.size() * 4);
    return ok();
}

int hiprt_bvh_validate(hiprt_host_scene s, hiprt_bvh b, int32_t max_tris) {
    if (!s || !b) return fail(HIPRT_ERR_ARG, "null argument");
    std::string e = rt::validate_bvh(s->sc, b->b, max_tris);
    if (!e.empty()) return fail(HIPRT_ERR_ARG, e);
    return ok();
}

void hiprt_bvh_free(hiprt_bvh b) { delete b; }

// --------------------------------------------------------------------- context

int hiprt_create(int32_t device, hiprt_context *out) {
    if (!out) return fail(HIPRT_ERR_ARG, "out is null");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(HIPRT_ERR_ARG, "no such device");
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<hiprt_context_s> c(new (std::nothrow) hiprt_context_s);
    if (!c) return fail(HIPRT_ERR_ARG, "out of host memory");
    c->device = device;
    if (!c->state(nullptr)) return fail(HIPRT_ERR_HIP, "cannot allocate the work counters");
    if (device >= kMaxDevices) return fail(HIPRT_ERR_UNSUPPORTED, "device index above 63");
    {
        // this device's copy of the allocation-free trace's scratch
        std::lock_guard<std::mutex> g(g_scratch_lock);
        if (!g_scratch_addr[device]) {
            void *ptr = nullptr;
            HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            if (!ptr) return fail(HIPRT_ERR_HIP, "no device address for the scratch array");
            g_scratch_addr[device] = ptr;
        }
    }
    *out = c.release();
    return ok();
}

int hiprt_destroy(hiprt_context ctx) {
    if (!ctx) return fail(HIPRT_ERR_ARG, "null context");
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return ok();
}

// ---------------------------------------------------------------- device scene

} // extern "C"

namespace hiprt_detail {

// Checks a flattened tree and computes its depth.  Children must have larger
// indices than their parent, which rules out cycles; leaf slot ranges must lie
// inside the triangle records.
std::string check_layout(rt::flat_scene &f, uint32_t &stack_need) {
    stack_need = 0;
    uint32_t n_slots = (uint32_t)f.tris.size();
    auto leaf_ok = [&](int32_t ref) {
        uint64_t s = rt::leaf_slot(ref), c = rt::leaf_count(ref);
        return s + c <= n_slots;
    };
    if (f.root_ref == rt::kInvalidRef) {
        if (!f.nodes2.empty() || !f.nodes4.empty()) return "nodes given for an empty tree";
        f.depth = 0;
        return "";
    }
    if (rt::ref_is_leaf(f.root_ref)) {
        if (!leaf_ok(f.root_ref)) return "root leaf slots out of range";
        f.depth = 0;
        return "";
    }
    size_t n = f.bvh_kind == rt::BVH2 ? f.nodes2.size() : f.nodes4.size();
    if (f.root_ref < 0 || (size_t)f.root_ref >= n) return "root index out of range";
    std::vector<uint32_t> depth(n, 1), need(n, 0);
    for (size_t i = n; i-- > 0;) {
        int32_t refs[4];
        int nref = 0;
        if (f.bvh_kind == rt::BVH2) {
            refs[0] = f.nodes2[i].ref_l;
            refs[1] = f.nodes2[i].ref_r;
            nref = 2;
        } else {
            for (int k = 0; k < 4; ++k) refs[k] = f.nodes4[i].ref[k];
            nref = 4;
        }
        int valid = 0;
        for (int k = 0; k < nref; ++k) {
            int32_t r = refs[k];
            if (r == rt::kInvalidRef) {
                if (f.bvh_kind == rt::BVH2) return "binary node with an absent child";
                continue;
            }
            ++valid;
            if (rt::ref_is_leaf(r)) {
                if (!leaf_ok(r)) return "leaf slots out of range";
                continue;
            }
            if ((size_t)r <= i || (size_t)r >= n) return "child index does not increase along an edge";
            depth[i] = std::max(depth[i], depth[r] + 1);
        }
        if (valid == 0) return "node without children";
        // a visit of node i pushes at most valid-1 entries, then descends
        uint32_t below = 0;
        for (int k = 0; k < nref; ++k)
            if (refs[k] != rt::kInvalidRef && !rt::ref_is_leaf(refs[k])) below = std::max(below, need[refs[k]]);
        need[i] = (uint32_t)(valid - 1) + below;
    }
    f.depth = depth[f.root_ref];
    stack_need = need[f.root_ref];
    return "";
}

void set_desc(hiprt_scene_s *s, uint32_t stack_need) {
    const rt::flat_scene &h = s->host;
    s->desc.nodes = s->nodes.p;
    s->desc.tris = static_cast<const rt::tri_rec *>(s->tris.p);
    s->desc.slot_ids = static_cast<const uint32_t *>(s->slot_ids.p);
    s->desc.root_ref = h.root_ref;
    s->desc.bvh = h.bvh_kind;
    s->desc.tri = h.tri_kind;
    s->desc.depth = (int32_t)h.depth;
    s->stack_need = (int)stack_need;
    {
        // box bounds of absent 4-wide children are +-inf, never NaN
        const float *fv = h.bvh_kind == rt::BVH2 ? reinterpret_cast<const float *>(h.nodes2.data())
                                                 : reinterpret_cast<const float *>(h.nodes4.data());
        size_t nf = h.bvh_kind == rt::BVH2 ? h.nodes2.size() * 16 : h.nodes4.size() * 32;
        size_t row = h.bvh_kind == rt::BVH2 ? 16 : 32, boxes = h.bvh_kind == rt::BVH2 ? 12 : 24;
        bool nan_free = true;
        for (size_t k = 0; k < nf && nan_free; ++k)
            if (k % row < boxes && std::isnan(fv[k])) nan_free = false;
        s->desc.nan_free = nan_free ? 1 : 0;
    }
}

} // namespace hiprt_detail

namespace {

using hiprt_detail::check_layout;

int upload_flat(hiprt_context ctx, rt::flat_scene &&f, hiprt_scene *out) {
    uint32_t stack_need = 0;
    std::string e = check_layout(f, stack_need);
    if (!e.empty()) return fail(HIPRT_ERR_ARG, "invalid BVH layout: " + e);
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<hiprt_scene_s> s(new (std::nothrow) hiprt_scene_s);
    if (!s) return fail(HIPRT_ERR_ARG, "out of host memory");
    s->ctx = ctx;
    s->host = std::move(f);
    const rt::flat_scene &h = s->host;
    if (h.bvh_kind == rt::BVH2) {
        HIP_TRY(s->nodes.upload(h.nodes2.data(), h.nodes2.size() * sizeof(rt::node2)));
    } else {
        HIP_TRY(s->nodes.upload(h.nodes4.data(), h.nodes4.size() * sizeof(rt::node4)));
    }
    HIP_TRY(s->tris.upload(h.tris.data(), h.tris.size() * sizeof(rt::tri_rec)));
    HIP_TRY(s->slot_ids.upload(h.slot_ids.data(), h.slot_ids.size() * 4));
    hiprt_detail::set_desc(s.get(), stack_need);
    *out = s.release();
    return ok();
}

} // namespace

extern "C" {

int hiprt_scene_commit(hiprt_context ctx, hiprt_host_scene s, hiprt_bvh b, int32_t bvh_format, int32_t tri_format,
                       hiprt_scene *out) {
    if (!ctx || !s || !b || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if (bvh_format != HIPRT_BVH2 && bvh_format != HIPRT_BVH4) return fail(HIPRT_ERR_ARG, "unknown bvh_format");
    if (tri_format != HIPRT_TRI_EXACT && tri_format != HIPRT_TRI_WOOP) return fail(HIPRT_ERR_ARG, "unknown tri_format");
    for (uint32_t t : b->b.index)
        if (t >= s->sc.n_tris()) return fail(HIPRT_ERR_ARG, "the BVH names a triangle the scene does not have");
    rt::flat_scene f;
    std::string e = rt::flatten(s->sc, b->b, bvh_format, tri_format, f);
    if (!e.empty()) return fail(HIPRT_ERR_UNSUPPORTED, e);
    hiprt_scene ds = nullptr;
    int st = upload_flat(ctx, std::move(f), &ds);
    if (st != HIPRT_OK) return st;
    std::unique_ptr<hiprt_scene_s> guard(ds);
    // Shading data: per triangle A.xyz + material id, B.xyz, C.xyz; per
    // material albedo and emission.
    const rt::scene &sc = s->sc;
    uint32_t nt = sc.n_tris();
    std::vector<float> geo(12ull * nt);
    for (uint32_t t = 0; t < nt; ++t) {
        rt::vec3f A = sc.a(t), B = sc.b(t), C = sc.c(t);
        uint32_t m = sc.mat.empty() ? 0u : sc.mat[t];
        float mb;
        std::memcpy(&mb, &m, 4);
        float v[12] = {A.x, A.y, A.z, mb, B.x, B.y, B.z, 0.0f, C.x, C.y, C.z, 0.0f};
        std::memcpy(&geo[12ull * t], v, sizeof v);
    }
    std::vector<float> mats(8 * (sc.mats.empty() ? 1 : sc.mats.size()), 0.0f);
    for (size_t m = 0; m < sc.mats.size(); ++m) {
        const rt::material &mt = sc.mats[m];
        float v[8] = {mt.albedo.x, mt.albedo.y, mt.albedo.z, 0.0f, mt.emissive.x, mt.emissive.y, mt.emissive.z, 0.0f};
        std::memcpy(&mats[8 * m], v, sizeof v);
    }
    for (uint32_t m : sc.mat)
        if (m >= sc.mats.size()) return fail(HIPRT_ERR_ARG, "triangle material id out of range");
    HIP_TRY(ds->tri_geo.upload(geo.data(), geo.size() * 4));
    HIP_TRY(ds->mats.upload(mats.data(), mats.size() * 4));
    HIP_TRY(ds->tri_vidx.upload(sc.idx.data(), sc.idx.size() * 4));
    // Shading normals, when the scene has them: the three corners' normals
    // per triangle (3 float4, w unused), so a hit reads one 48-byte row.
    if (!sc.norm.empty() && sc.norm.size() == 3 * sc.pos.size() && nt) {
        std::vector<float> tn(12ull * nt, 0.0f);
        for (uint32_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) tn[12ull * t + 4 * k + c] = sc.norm[3ull * sc.idx[3ull * t + k] + c];
        HIP_TRY(ds->tri_ns.upload(tn.data(), tn.size() * 4));
    }
    // Opacity data for the alpha-aware tracers, when a material has a texture.
    bool textured = false;
    for (const rt::material &m : sc.mats) textured = textured || (m.tex >= 0 && (size_t)m.tex < sc.texs.size());
    if (textured && nt) {
        std::vector<float> atri(8ull * nt, 0.0f);
        for (uint32_t t = 0; t < nt; ++t) {
            uint32_t m = sc.mat.empty() ? 0u : sc.mat[t];
            int32_t tex = m < sc.mats.size() ? sc.mats[m].tex : -1;
            uint32_t texb = tex >= 0 ? (uint32_t)tex : 0xffffffffu;
            float *o = &atri[8ull * t];
            for (int k = 0; k < 3; ++k) {
                uint32_t v = sc.idx[3ull * t + k];
                bool have = sc.tc.size() >= 2ull * (v + 1);
                float uu = have ? sc.tc[2ull * v] : 0.0f, vv = have ? sc.tc[2ull * v + 1] : 0.0f;
                o[2 * k] = uu;   // a.u a.v b.u b.v | c.u c.v
                o[2 * k + 1] = vv;
            }
            std::memcpy(&o[6], &texb, 4);
        }
        std::vector<uint32_t> atex(4 * sc.texs.size(), 0u);
        std::vector<uint8_t> mask;
        for (size_t k = 0; k < sc.texs.size(); ++k) {
            const rt::texture &tx = sc.texs[k];
            atex[4 * k] = (uint32_t)mask.size();
            atex[4 * k + 1] = tx.w;
            atex[4 * k + 2] = tx.h;
            for (size_t i = 0; i < (size_t)tx.w * tx.h; ++i) mask.push_back(tx.rgba[4 * i + 3] > 0.5f ? 1 : 0);
        }
        if (mask.size() > 0xffffffffull) return fail(HIPRT_ERR_ARG, "textures too large");
        HIP_TRY(ds->alpha_tri.upload(atri.data(), atri.size() * 4));
        HIP_TRY(ds->alpha_tex.upload(atex.data(), atex.size() * 4));
        HIP_TRY(ds->alpha_mask.upload(mask.data(), mask.size()));
        ds->desc.alpha_tri = ds->alpha_tri.p;
        ds->desc.alpha_tex = ds->alpha_tex.p;
        ds->desc.alpha_mask = ds->alpha_mask.as<const uint8_t>();
    }
    ds->n_tris = nt;
    ds->n_vert = (uint32_t)sc.pos.size();
    *out = guard.release();
    return ok();
}

int hiprt_scene_upload(hiprt_context ctx, const void *nodes, uint32_t n_nodes, int32_t bvh_format,
                       const void *tri_records, const uint32_t *slot_ids, uint32_t n_slots, int32_t tri_format,
                       int32_t root_ref, hiprt_scene *out) {
    if (!ctx || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if ((n_nodes && !nodes) || (n_slots && (!tri_records || !slot_ids))) return fail(HIPRT_ERR_ARG, "null array");
    if (bvh_format != HIPRT_BVH2 && bvh_format != HIPRT_BVH4) return fail(HIPRT_ERR_ARG, "unknown bvh_format");
    if (tri_format != HIPRT_TRI_EXACT && tri_format != HIPRT_TRI_WOOP) return fail(HIPRT_ERR_ARG, "unknown tri_format");
    rt::flat_scene f;
    f.bvh_kind = bvh_format;
    f.tri_kind = tri_format;
    f.root_ref = root_ref;
    if (bvh_format == HIPRT_BVH2) {
        const auto *p = static_cast<const rt::node2 *>(nodes);
        f.nodes2.assign(p, p + n_nodes);
    } else {
        const auto *p = static_cast<const rt::node4 *>(nodes);
        f.nodes4.assign(p, p + n_nodes);
    }
    const auto *t = static_cast<const rt::tri_rec *>(tri_records);
    f.tris.assign(t, t + n_slots);
    f.slot_ids.assign(slot_ids, slot_ids + n_slots);
    return upload_flat(ctx, std::move(f), out);
}

int hiprt_scene_layout(hiprt_scene sc, uint32_t *n_nodes, uint32_t *n_slots, int32_t *root_ref, uint32_t *depth,
                       int32_t *bvh_format, int32_t *tri_format) {
    if (!sc) return fail(HIPRT_ERR_ARG, "null scene");
    const rt::flat_scene &h = sc->host;
    if (n_nodes) *n_nodes = (uint32_t)(h.bvh_kind == rt::BVH2 ? h.nodes2.size() : h.nodes4.size());
    if (n_slots) *n_slots = (uint32_t)h.tris.size();
    if (root_ref) *root_ref = h.root_ref;
    if (depth) *depth = h.depth;
    if (bvh_format) *bvh_format = h.bvh_kind;
    if (tri_format) *tri_format = h.tri_kind;
    return ok();
}

int hiprt_scene_download(hiprt_scene sc, void *nodes, void *tri_records, uint32_t *slot_ids) {
    if (!sc) return fail(HIPRT_ERR_ARG, "null scene");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    // sizes of the layout (hiprt_scene_layout), not of the allocations, which
    // a device build makes for its worst case
    const rt::flat_scene &h = sc->host;
    size_t node_bytes = h.bvh_kind == rt::BVH2 ? h.nodes2.size() * sizeof(rt::node2) : h.nodes4.size() * sizeof(rt::node4);
    size_t rec_bytes = h.tris.size() * sizeof(rt::tri_rec), id_bytes = h.slot_ids.size() * 4;
    if (nodes && node_bytes) HIP_TRY(hipMemcpy(nodes, sc->nodes.p, node_bytes, hipMemcpyDeviceToHost));
    if (tri_records && rec_bytes) HIP_TRY(hipMemcpy(tri_records, sc->tris.p, rec_bytes, hipMemcpyDeviceToHost));
    if (slot_ids && id_bytes) HIP_TRY(hipMemcpy(slot_ids, sc->slot_ids.p, id_bytes, hipMemcpyDeviceToHost));
    return ok();
}

int hiprt_scene_refit(hiprt_scene sc, const float *d_positions, void *stream) {
    if (!sc || !d_positions) return fail(HIPRT_ERR_ARG, "null argument");
    if (!sc->tri_vidx.p) {
        if (sc->host.tris.empty()) return ok();
        return fail(HIPRT_ERR_UNSUPPORTED, "refit needs a scene committed from a host scene");
    }
    HIP_TRY(hipSetDevice(sc->ctx->device));
    const rt::flat_scene &h = sc->host;
    if (sc->refit_off.empty()) {
        // inner nodes grouped by depth, deepest level first
        std::vector<std::vector<int32_t>> levels;
        if (h.root_ref >= 0) {
            std::vector<int32_t> cur{h.root_ref};
            while (!cur.empty()) {
                levels.push_back(cur);
                std::vector<int32_t> next;
                for (int32_t id : cur) {
                    int32_t refs[4] = {rt::kInvalidRef, rt::kInvalidRef, rt::kInvalidRef, rt::kInvalidRef};
                    if (h.bvh_kind == rt::BVH2) {
                        refs[0] = h.nodes2[id].ref_l;
                        refs[1] = h.nodes2[id].ref_r;
                    } else {
                        for (int c = 0; c < 4; ++c) refs[c] = h.nodes4[id].ref[c];
                    }
                    for (int32_t r : refs)
                        if (r >= 0 && r != rt::kInvalidRef) next.push_back(r);
                }
                cur.swap(next);
            }
        }
        std::vector<int32_t> list;
        sc->refit_off.push_back(0);
        for (auto it = levels.rbegin(); it != levels.rend(); ++it) {
            list.insert(list.end(), it->begin(), it->end());
            sc->refit_off.push_back((uint32_t)list.size());
        }
        HIP_TRY(sc->refit_list.upload(list.data(), list.size() * 4));
    }
    rt::dev::refit_args a{};
    a.nodes = sc->nodes.p;
    a.bvh = h.bvh_kind;
    a.tris = sc->tris.as<rt::tri_rec>();
    a.slot_ids = sc->slot_ids.as<uint32_t>();
    a.tri = h.tri_kind;
    a.n_slots = (uint32_t)h.tris.size();
    a.tri_geo = sc->tri_geo.as<float>();
    a.n_tris = sc->n_tris;
    a.pos = d_positions;
    a.vidx = sc->tri_vidx.as<uint32_t>();
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(rt::dev::launch_refit_records(a, s));
    const int32_t *list = sc->refit_list.as<int32_t>();
    for (size_t k = 0; k + 1 < sc->refit_off.size(); ++k)
        HIP_TRY(rt::dev::launch_refit_level(a, list + sc->refit_off[k], sc->refit_off[k + 1] - sc->refit_off[k], s));
    sc->desc.nan_free = 1;   // select-based min/max from an empty box never store NaN
    return ok();
}

int hiprt_scene_stack_need(hiprt_scene sc, uint32_t *stack_need) {
    if (!sc || !stack_need) return fail(HIPRT_ERR_ARG, "null argument");
    *stack_need = (uint32_t)sc->stack_need;
    return ok();
}

int hiprt_scene_destroy(hiprt_scene sc) {
    if (!sc) return fail(HIPRT_ERR_ARG, "null scene");
    (void)hipSetDevice(sc->ctx->device);
    delete sc;
    return ok();
}

// ----------------------------------------------------------------------- trace

static const char *const kTracerNames[2][HIPRT_TRACER_COUNT] = {
    {"while-while", "persistent-while-while", "persistent-speculative-while-while", "dynamic-while-while"},
    {"while-while-alpha", "persistent-while-while-alpha", "persistent-speculative-while-while-alpha",
     "dynamic-while-while-alpha"}};

// The reference's other CUDA variants (rt/cuda/platform.cpp:107-122) are
// loop structures of the same near-first traversal: each name selects the
// kernel here that runs that traversal.
static const struct {
    const char *name;
    int id;
} kTracerAliases[] = {
    {"simple", HIPRT_TRACER_WHILE_WHILE},
    {"if-if", HIPRT_TRACER_WHILE_WHILE},
    {"persistent-if-if", HIPRT_TRACER_PERSISTENT_WHILE_WHILE},
    {"speculative-while-while", HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE},
    // the reference's links (rt/cuda/platform.cpp:55-56): the deterministic
    // while-while traversal
    {"default", HIPRT_TRACER_WHILE_WHILE},
    {"find closest hits", HIPRT_TRACER_WHILE_WHILE},
};

int hiprt_tracer_id(const char *name) {
    if (!name) return -1;
    for (int a = 0; a < 2; ++a)
        for (int i = 0; i < HIPRT_TRACER_COUNT; ++i)
            if (std::strcmp(name, kTracerNames[a][i]) == 0) return i | (a ? HIPRT_TRACER_ALPHA : 0);
    const size_t len = std::strlen(name);
    const bool alpha = len > 6 && std::strcmp(name + len - 6, "-alpha") == 0;
    for (const auto &al : kTracerAliases) {
        const size_t n = std::strlen(al.name);
        if ((alpha ? len == n + 6 : len == n) && std::strncmp(name, al.name, n) == 0)
            return al.id | (alpha ? HIPRT_TRACER_ALPHA : 0);
    }
    return -1;
}

const char *hiprt_tracer_name(int tracer) {
    int a = (tracer & HIPRT_TRACER_ALPHA) ? 1 : 0;
    int t = tracer & ~HIPRT_TRACER_ALPHA;
    if (t < 0 || t >= HIPRT_TRACER_COUNT) return "";
    return kTracerNames[a][t];
}

namespace {

// noalloc: work counters and spill column from this device's static
// scratch instead of the stream state's buffers (allocates nothing).
int trace_common(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, uint8_t *occluded, const hiprt_hits *hits,
                 bool any, int32_t tracer, const hiprt_tracer_opts *opts, void *stream, bool noalloc = false) {
    if (!sc || !rays) return fail(HIPRT_ERR_ARG, "null argument");
    const bool alpha = (tracer & HIPRT_TRACER_ALPHA) != 0;
    tracer &= ~HIPRT_TRACER_ALPHA;
    if (tracer < 0 || tracer >= HIPRT_TRACER_COUNT) return fail(HIPRT_ERR_ARG, "unknown tracer id");
    if (n > 0xF0000000u) return fail(HIPRT_ERR_ARG, "too many rays in one batch");
    if (n == 0) return ok();
    if (!rays->d_ox || !rays->d_oy || !rays->d_oz || !rays->d_tmin || !rays->d_dx || !rays->d_dy || !rays->d_dz ||
        !rays->d_tmax)
        return fail(HIPRT_ERR_ARG, "null ray plane");
    if (!any && (!hits || !hits->d_t || !hits->d_u || !hits->d_v || !hits->d_id))
        return fail(HIPRT_ERR_ARG, "null hit plane");
    if (any && !occluded && !(hits && hits->d_t))
        return fail(HIPRT_ERR_ARG, "any-hit needs an occlusion array or hit planes");
    if (any && hits && hits->d_t && (!hits->d_u || !hits->d_v || !hits->d_id))
        return fail(HIPRT_ERR_ARG, "null hit plane");
    HIP_TRY(hipSetDevice(sc->ctx->device));
    rt::ray_soa r{rays->d_ox, rays->d_oy, rays->d_oz, rays->d_tmin, rays->d_dx, rays->d_dy, rays->d_dz, rays->d_tmax};
    rt::hit_soa h{nullptr, nullptr, nullptr, nullptr};
    if (hits) h = rt::hit_soa{hits->d_t, hits->d_u, hits->d_v, hits->d_id};
    hipStream_t s = static_cast<hipStream_t>(stream);
    const rt::scene_desc desc = hiprt_detail::desc_for(sc, alpha);
    int need = sc->stack_need;
    int stack = (opts && opts->stack_config > 0 && opts->stack_config < rt::dev::kStackConfigs) ? opts->stack_config : 0;
    if (desc.bvh == rt::BVH4 && tracer == HIPRT_TRACER_WHILE_WHILE)
        return fail(HIPRT_ERR_UNSUPPORTED, "4-wide BVH needs a persistent tracer");
    if (tracer == HIPRT_TRACER_WHILE_WHILE) {
        if (need > rt::dev::kPlainStackCap - 1)
            return fail(HIPRT_ERR_UNSUPPORTED, "BVH deeper than the traversal stack (" + std::to_string(need) + " entries)");
        HIP_TRY(rt::dev::launch_trace_plain(desc, r, h, occluded, n, any, s,
                                            opts ? reinterpret_cast<unsigned long long *>(opts->d_fetch_stats)
                                                 : nullptr));
        return ok();
    }
    rt::dev::persistent_cfg cfg{};
    cfg.speculative = tracer >= HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE;
    cfg.refill = tracer == HIPRT_TRACER_DYNAMIC_WHILE_WHILE;
    cfg.refill_thr = (opts && opts->refill_threshold > 0) ? opts->refill_threshold : 40;
    if (cfg.refill_thr > 64) cfg.refill_thr = 64;
    cfg.blocks = (opts && opts->blocks > 0) ? opts->blocks : 0;
    cfg.stats = opts ? reinterpret_cast<unsigned long long *>(opts->d_fetch_stats) : nullptr;
    cfg.stack = stack;
    cfg.chunks = (opts && opts->work_chunks == 1) ? 1 : 8;
    cfg.park = (opts && opts->park_leaves > 0 && opts->park_leaves < rt::dev::kParkClosest) ? opts->park_leaves
                                                                                          : rt::dev::kParkClosest;
    uint32_t *counter = nullptr;
    if (noalloc) {
        char *base = nullptr;
        {
            std::lock_guard<std::mutex> g(g_scratch_lock);
            base = static_cast<char *>(g_scratch_addr[sc->ctx->device]);
        }
        if (!base) return fail(HIPRT_ERR_HIP, "the device's scratch is not set up");
        counter = reinterpret_cast<uint32_t *>(base);
        cfg.spill_cap = rt::dev::persistent_spill_entries(desc.bvh, cfg.stack, need);
        if (opts && opts->spill_limit > 0 && cfg.spill_cap > (uint32_t)opts->spill_limit)
            cfg.spill_cap = (uint32_t)opts->spill_limit;
        if (cfg.spill_cap) {
            int blocks = 0;
            HIP_TRY(rt::dev::persistent_grid(desc, any, cfg, &blocks));
            const size_t bytes = rt::dev::persistent_spill_bytes(blocks, cfg.spill_cap);
            if (bytes > kNoAllocSpillBytes)
                return fail(HIPRT_ERR_UNSUPPORTED,
                            "the traversal stack's spill column needs " + std::to_string(bytes) + " bytes (" +
                                std::to_string(blocks) + " blocks x " + std::to_string(cfg.spill_cap) +
                                " entries per lane); the allocation-free trace holds " +
                                std::to_string(kNoAllocSpillBytes) +
                                ": use fewer blocks, a stack configuration with more entries, or hiprt_trace_closest");
            cfg.spill = reinterpret_cast<int32_t *>(base + rt::dev::kCounterBytes);
            cfg.spill_bytes = kNoAllocSpillBytes;
        }
    } else {
        hiprt_stream_state *st = sc->ctx->state(stream);
        if (!st) return fail(HIPRT_ERR_HIP, "cannot allocate the stream's work counters");
        HIP_TRY(hiprt_detail::attach_spill(st, s, desc, any, false, need, cfg, opts ? opts->spill_limit : 0));
        counter = st->counter.as<uint32_t>();
    }
    HIP_TRY(rt::dev::launch_trace_persistent(desc, r, h, occluded, n, nullptr, counter, any, cfg, s));
    return ok();
}

} // namespace

int hiprt_trace_closest(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, const hiprt_hits *hits, int32_t tracer,
                        const hiprt_tracer_opts *opts, void *stream) {
    return trace_common(sc, rays, n, nullptr, hits, false, tracer, opts, stream);
}

int hiprt_trace_any(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, uint8_t *d_occluded, const hiprt_hits *hits,
                    int32_t tracer, const hiprt_tracer_opts *opts, void *stream) {
    return trace_common(sc, rays, n, d_occluded, hits, true, tracer, opts, stream);
}

int hiprt_trace_closest_noalloc(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, const hiprt_hits *hits,
                                int32_t tracer, const hiprt_tracer_opts *opts, void *stream) {
    return trace_common(sc, rays, n, nullptr, hits, false, tracer, opts, stream, true);
}

} // extern "C"
