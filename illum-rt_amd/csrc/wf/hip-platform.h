// The "hip" wavefront platform: RTGI's platform/tracer/step interface
// (wavefront-rt.h) implemented on top of the C ABI of libillum_hiprt.so
// only.  It plays the role of rt/cuda/platform.cpp and rt/cuda/base.{h,cpp}
// for the MI355X: tracers are registered under names, selected by
// "raytracer <name>", the BVH is built on the host by "build accel struct"
// with the GPU defaults (SAH, 16 planes, 4 triangles per leaf, ESC) and
// flattened into device memory, rays and hits live in device SoA buffers.
#pragma once

#include "wavefront-rt.h"

#include "../../../include/illum_hiprt.h"

#include <iosfwd>
#include <memory>
#include <string>
#include <vector>

namespace wf {

struct pointlight {
    float pos[3], col[3];
};

// The scene as the driver describes it: geometry, camera, lights.
struct scene {
    hiprt_host_scene geometry = nullptr;
    float cam_pos[3] = {0, 0, 0}, cam_dir[3] = {0, 0, -1}, cam_up[3] = {0, 1, 0};
    float fovy = 65.0f;   // libgi/scene.h:164
    std::vector<pointlight> lights;
    ~scene() {
        if (geometry) hiprt_scene_free(geometry);
    }
};

// The parts of RTGI's render context this path needs (libgi/context.h).
struct render_context {
    int w = 1280, h = 720;
    int sppx = 1;
    std::vector<float> framebuffer;   // rgba per pixel, after "download framebuffer"
    wf::platform *platform = nullptr;
    wf::scene scene;
    void camera(float *cam12) const;   // pos, dir, up, fovy, w, h
};
extern render_context *rc;

namespace hip {

class platform;
extern platform *pf;

// Errors from the C ABI become std::runtime_error, as CHECK_CUDA_ERROR does
// (rt/cuda/cuda-helpers.cpp:5-12).
void check(int status, const char *what);

// Device memory owned by the platform (global_memory_buffer, rt/cuda/base.h:100-195).
struct device_buffer {
    void *p = nullptr;
    size_t bytes = 0;
    device_buffer() = default;
    device_buffer(const device_buffer &) = delete;
    device_buffer &operator=(const device_buffer &) = delete;
    ~device_buffer();
    void resize(size_t n);
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Per-pixel rays (8 planes), hits (4 planes) and occlusion bytes.
struct raydata : public wf::raydata {
    int w = 0, h = 0;
    device_buffer rays, hits, occluded;
    raydata(int w, int h);
    void resize(int w, int h);
    uint32_t n() const { return (uint32_t)w * (uint32_t)h; }
    hiprt_rays in() const;
    hiprt_ray_buffers out() const;
    hiprt_hits hit() const;
};

// A registered tracer: one kernel variant of the library.
struct batch_rt : public wf::batch_ray_tracer {
    int tracer;
    raydata *rd = nullptr;
    // "incoherence r1 r2" (rt/cuda/base.cpp:120-133, base.h:405-407): the
    // camera step then traces rays between two balls instead
    bool use_incoherence = false;
    float incoherence_r1 = 0, incoherence_r2 = 0;
    explicit batch_rt(int id) : tracer(id) {}
    void compute_closest_hit() override;
    void compute_any_hit() override;
    void use(wf::raydata *r) override { rd = dynamic_cast<raydata *>(r); }
};

class platform : public wf::platform {
public:
    hiprt_context ctx = nullptr;
    hiprt_bvh_params bvh_params{2, 4, 16, 1};
    int tri_format = HIPRT_TRI_EXACT;
    // options of the persistent tracers, set by "tracer-option <name> <value>"
    hiprt_tracer_opts opts{40, 0, nullptr, 0, 8, 0, 0};
    int ray_order = 0;   // path tracer bounce queues: 0 sorted, 1 shading order
    int samples = 1;     // path tracer: samples per pixel traced together (divides sppx)
    int regen_pool = 0;  // path tracer: path slots per pixel of path regeneration (0: off)
    hiprt_bvh bvh = nullptr;
    hiprt_scene ds = nullptr;          // device scene of the last commit
    wf::scene *sd = nullptr;
    batch_rt *rt = nullptr;            // the tracer steps use
    device_buffer framebuffer;         // float4 per pixel

    explicit platform(const std::vector<std::string> &args);
    ~platform() override;
    raydata *allocate_raydata() override;
    void commit_scene(wf::scene *s) override;
    bool interprete(const std::string &command, std::istringstream &in) override;
    void build();                      // host BVH + flatten + upload
    // scene steps "rotate scene [keeping original data]" / "drop scene view"
    void rotate_scene(bool keep);
    void drop_scene_view();
    std::vector<std::vector<float>> views;   // positions kept by "rotate scene keeping original data"
};

// Algorithms (the gi/ step graphs): primary-wf (gi/primary-hit.cpp:85-117),
// direct-wf with its uniform, cosine and light sampling modes
// (gi/direct.cpp:470-585) and pt-wf, the wavefront path tracer of this build.
std::unique_ptr<simple_algorithm> make_algorithm(const std::string &name);

// Step times collected by the platform timer ("stats print" / "stats clear").
void print_stats(std::ostream &out);
void clear_stats();

} // namespace hip
} // namespace wf
