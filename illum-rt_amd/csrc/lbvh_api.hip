// C ABI of the device BVH build: triangles in device memory in, a committed
// scene out, without a host-side builder.
#include "capi_internal.h"

#include "device/lbvh_kernels.h"

#include <cstring>
#include <memory>
#include <new>
#include <vector>

using hiprt_detail::dbuf;
using hiprt_detail::fail;
using hiprt_detail::ok;

namespace {

// Shading copy of the vertices (A.xyz + material id bits, B.xyz, C.xyz per
// triangle), as hiprt_scene_commit builds it on the host.
__global__ __launch_bounds__(256) void k_tri_geo(const float *pos, const uint32_t *vidx, const uint32_t *mat,
                                                 uint32_t n_vert, uint32_t n_tris, float *geo) {
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n_tris) return;
    float *g = geo + 12ull * t;
    for (int c = 0; c < 3; ++c) {
        uint32_t v = vidx[3ull * t + c];
        if (v >= n_vert) v = 0;   // reported by the build
        g[4 * c] = pos[3ull * v];
        g[4 * c + 1] = pos[3ull * v + 1];
        g[4 * c + 2] = pos[3ull * v + 2];
        g[4 * c + 3] = 0.0f;
    }
    g[3] = __uint_as_float(mat ? mat[t] : 0u);
}

__global__ void k_max_u32(const uint32_t *p, uint32_t n, uint32_t *out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) atomicMax(out, p[i]);
}

} // namespace

extern "C" {

int hiprt_scene_build(hiprt_context ctx, const float *d_positions, uint32_t n_vert, const uint32_t *d_indices,
                      uint32_t n_tris, const uint32_t *d_materials, const float *materials, uint32_t n_materials,
                      int32_t max_leaf, int32_t tri_format, hiprt_scene *out, float *device_ms, void *stream) {
    if (!ctx || !out) return fail(HIPRT_ERR_ARG, "null argument");
    if ((n_tris && (!d_positions || !d_indices)) || (n_materials && !materials))
        return fail(HIPRT_ERR_ARG, "null array");
    if (max_leaf < 1 || max_leaf > (int32_t)rt::kMaxLeafCount) return fail(HIPRT_ERR_ARG, "max_leaf must be 1..32");
    if (tri_format != HIPRT_TRI_EXACT && tri_format != HIPRT_TRI_WOOP) return fail(HIPRT_ERR_ARG, "unknown tri_format");
    if (n_tris > rt::kMaxLeafSlot) return fail(HIPRT_ERR_ARG, "too many triangles for the 26-bit slot field");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::unique_ptr<hiprt_scene_s> sc(new (std::nothrow) hiprt_scene_s);
    if (!sc) return fail(HIPRT_ERR_ARG, "out of host memory");
    sc->ctx = ctx;
    const uint32_t n = n_tris, m = n > 1 ? n - 1 : 1;
    dbuf scratch, small;
    HIP_TRY(sc->nodes.alloc(sizeof(rt::node2) * m));
    HIP_TRY(sc->tris.alloc(sizeof(rt::tri_rec) * (n ? n : 1)));
    HIP_TRY(sc->slot_ids.alloc(4ull * (n ? n : 1)));
    HIP_TRY(scratch.alloc(rt::dev::lbvh_scratch_bytes(n)));
    HIP_TRY(small.alloc(64));
    int32_t *d_root = small.as<int32_t>();
    uint32_t *d_counts = reinterpret_cast<uint32_t *>(d_root + 4);

    rt::dev::lbvh_in in{d_positions, d_indices, n_vert, n, max_leaf, tri_format};
    rt::dev::lbvh_out o{sc->nodes.as<rt::node2>(), sc->tris.as<rt::tri_rec>(), sc->slot_ids.as<uint32_t>(), d_root,
                        d_counts};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (device_ms) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, s));
    }
    hipError_t le = rt::dev::launch_lbvh(in, o, scratch.p, s);
    if (device_ms && le == hipSuccess) le = hipEventRecord(e1, s);
    if (le == hipSuccess) le = hipStreamSynchronize(s);
    if (le == hipSuccess && device_ms) le = hipEventElapsedTime(device_ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (le != hipSuccess) return fail(HIPRT_ERR_HIP, std::string("device BVH build: ") + hipGetErrorString(le));

    int32_t root = 0;
    uint32_t counts[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpy(&root, d_root, 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(counts, d_counts, 16, hipMemcpyDeviceToHost));
    if (counts[2]) return fail(HIPRT_ERR_ARG, std::to_string(counts[2]) + " triangles name a vertex past n_vert");
    if (counts[3]) return fail(HIPRT_ERR_HIP, "device BVH build produced a node off the tree");
    if (counts[0] > m) return fail(HIPRT_ERR_HIP, "device BVH build: node count out of range");
    if (d_materials && n_tris) {
        dbuf mx;
        HIP_TRY(mx.alloc(4));
        HIP_TRY(hipMemset(mx.p, 0, 4));
        hipLaunchKernelGGL(k_max_u32, dim3((n + 255) / 256), dim3(256), 0, s, d_materials, n, mx.as<uint32_t>());
        HIP_TRY(hipGetLastError());
        uint32_t mmax = 0;
        HIP_TRY(hipMemcpy(&mmax, mx.p, 4, hipMemcpyDeviceToHost));
        if (mmax >= (n_materials ? n_materials : 1)) return fail(HIPRT_ERR_ARG, "triangle material id out of range");
    }

    // Host copy of the flattened tree: the same checks as uploaded trees pass
    // (children after their parent, slot ranges, stack need), and the refit
    // and download entry points work on it.
    rt::flat_scene &h = sc->host;
    h.bvh_kind = rt::BVH2;
    h.tri_kind = tri_format;
    h.root_ref = root;
    h.nodes2.resize(counts[0]);
    h.tris.resize(n);
    h.slot_ids.resize(n);
    if (counts[0]) HIP_TRY(hipMemcpy(h.nodes2.data(), sc->nodes.p, sizeof(rt::node2) * counts[0], hipMemcpyDeviceToHost));
    if (n) {
        HIP_TRY(hipMemcpy(h.tris.data(), sc->tris.p, sizeof(rt::tri_rec) * n, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(h.slot_ids.data(), sc->slot_ids.p, 4ull * n, hipMemcpyDeviceToHost));
    }
    uint32_t stack_need = 0;
    std::string err = hiprt_detail::check_layout(h, stack_need);
    if (!err.empty()) return fail(HIPRT_ERR_HIP, "device BVH build produced an invalid tree: " + err);
    if (h.depth != counts[1]) return fail(HIPRT_ERR_HIP, "device BVH build: depth mismatch");
    hiprt_detail::set_desc(sc.get(), stack_need);

    // shading data and vertex indices, as hiprt_scene_commit provides them
    HIP_TRY(sc->tri_geo.alloc(48ull * (n ? n : 1)));
    if (n) {
        hipLaunchKernelGGL(k_tri_geo, dim3((n + 255) / 256), dim3(256), 0, s, d_positions, d_indices, d_materials,
                           n_vert, n, sc->tri_geo.as<float>());
        HIP_TRY(hipGetLastError());
    }
    std::vector<float> mats(8 * (n_materials ? n_materials : 1), 0.0f);
    if (n_materials) std::memcpy(mats.data(), materials, 32ull * n_materials);
    HIP_TRY(sc->mats.upload(mats.data(), mats.size() * 4));
    HIP_TRY(sc->tri_vidx.alloc(12ull * (n ? n : 1)));
    if (n) HIP_TRY(hipMemcpyAsync(sc->tri_vidx.p, d_indices, 12ull * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    sc->n_tris = n;
    sc->n_vert = n_vert;
    *out = sc.release();
    return ok();
}

} // extern "C"
