// Internal state behind the C ABI's opaque handles, shared by the .hip files
// that implement include/illum_hiprt.h.
#pragma once

#include "../../include/illum_hiprt.h"

#include "core/layout.h"
#include "device/kernels.h"
#include "host/bvh.h"
#include "host/flatten.h"
#include "host/scene.h"

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace hiprt_detail {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string &msg);
inline int ok() { return HIPRT_OK; }

#define HIP_TRY(call)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            return ::hiprt_detail::fail(HIPRT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// Owned device allocation.
struct dbuf {
    void *p = nullptr;
    size_t bytes = 0;
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        release();
        bytes = n;
        return n ? hipMalloc(&p, n) : hipSuccess;
    }
    hipError_t upload(const void *src, size_t n) {
        hipError_t e = alloc(n);
        if (e != hipSuccess || n == 0) return e;
        return hipMemcpy(p, src, n, hipMemcpyHostToDevice);
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

} // namespace hiprt_detail

struct hiprt_host_scene_s {
    rt::scene sc;
};

struct hiprt_bvh_s {
    rt::bvh b;
    rt::bvh_stats st;
};

// Per-stream state of a context: launches on one stream are ordered, so
// they can share a work-counter block and a descriptor table; launches on
// different streams get their own.
struct hiprt_stream_state {
    hiprt_detail::dbuf counter;       // work counters of the persistent kernels
    hiprt_detail::dbuf batch_table;   // descriptor table of the last batched trace on the stream
    hiprt_detail::dbuf spill;         // stack spill buffer of the persistent kernels (grown on demand)
    hiprt_detail::dbuf adjoint;       // sort scratch of the adjoint's vertex accumulation (grown on demand)
};

struct hiprt_context_s {
    int device = 0;
    std::mutex lock;                  // guards `streams`
    std::map<void *, std::unique_ptr<hiprt_stream_state>> streams;
    // The state of `stream` (created on first use, counters zeroed); null on
    // allocation failure.
    hiprt_stream_state *state(void *stream);
};

struct hiprt_scene_s {
    hiprt_context ctx = nullptr;
    rt::flat_scene host;                         // host copy of the uploaded layout
    hiprt_detail::dbuf nodes, tris, slot_ids;    // traversal data
    hiprt_detail::dbuf tri_geo, mats;            // shading data (scenes committed from a host scene)
    hiprt_detail::dbuf tri_vidx;                 // vertex indices per triangle (adjoint scatter)
    hiprt_detail::dbuf tri_ns;                   // corner shading normals per triangle (scenes with normals)
    hiprt_detail::dbuf alpha_tri, alpha_tex, alpha_mask;   // opacity data (scenes with textures)
    uint32_t n_tris = 0, n_vert = 0;
    int stack_need = 0;                          // most stack entries any traversal can hold
    hiprt_detail::dbuf light_rec, light_f, light_cdf;   // light list of the direct-lighting steps
    uint32_t n_lights = 0;
    float light_integral = 0.0f;
    hiprt_detail::dbuf refit_list;               // inner nodes by level, deepest first (refit)
    std::vector<uint32_t> refit_off;             // level boundaries in refit_list
    rt::scene_desc desc{};
};

namespace hiprt_detail {

// Checks a flattened tree (children after their parent, leaf slots in range,
// depth and node limits) and computes the most stack entries a traversal can
// hold.  Returns "" when the tree is valid.
std::string check_layout(rt::flat_scene &f, uint32_t &stack_need);

// Fills the scene descriptor from the device buffers and the host copy.
void set_desc(hiprt_scene_s *s, uint32_t stack_need);

// The scene's descriptor for one trace: the opacity test on when `alpha`
// is asked for and the scene has opacity data.
rt::scene_desc desc_for(const hiprt_scene_s *sc, bool alpha);

// Gives `cfg` the spill buffer a persistent launch on stream `s` needs for
// trees whose traversal can hold `need` stack entries (none when the
// configuration's VGPR and LDS entries hold them): the stream state's
// buffer, grown once the stream's earlier launches, which may still use the
// old one, have finished.  `multi` selects the descriptor-table kernel;
// spill_limit > 0 caps the entries per lane (hiprt_tracer_opts::spill_limit).
hipError_t attach_spill(hiprt_stream_state *st, hipStream_t s, const rt::scene_desc &d, bool any, bool multi,
                        int need, rt::dev::persistent_cfg &cfg, int spill_limit = 0);

} // namespace hiprt_detail
