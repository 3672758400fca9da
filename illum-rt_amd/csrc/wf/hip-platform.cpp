#include "hip-platform.h"

#include <hip/hip_runtime_api.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <iostream>
#include <map>
#include <vector>

namespace wf {

render_context *rc = nullptr;

void render_context::camera(float *cam) const {
    for (int k = 0; k < 3; ++k) {
        cam[k] = scene.cam_pos[k];
        cam[3 + k] = scene.cam_dir[k];
        cam[6 + k] = scene.cam_up[k];
    }
    cam[9] = scene.fovy;
    cam[10] = (float)w;
    cam[11] = (float)h;
}

namespace hip {

platform *pf = nullptr;

void check(int status, const char *what) {
    if (status != HIPRT_OK) throw std::runtime_error(std::string(what) + ": " + hiprt_last_error());
}

static void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

device_buffer::~device_buffer() {
    if (p) (void)hipFree(p);
}

void device_buffer::resize(size_t n) {
    if (n == bytes) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = n;
    if (n) hip_check(hipMalloc(&p, n), "hipMalloc");
}

// ---------------------------------------------------------------- timer

// HIP-event pairs per step name, folded into totals at synchronize()
// (rt/cuda/base.cpp:17-45).
struct event_timer : public wf::timer {
    std::map<std::string, std::pair<hipEvent_t, hipEvent_t>> events;
    static std::map<std::string, std::pair<double, long>> &totals() {
        static std::map<std::string, std::pair<double, long>> t;
        return t;
    }
    void start(const std::string &name) override {
        auto it = events.find(name);
        if (it == events.end()) {
            hipEvent_t a, b;
            hip_check(hipEventCreate(&a), "hipEventCreate");
            hip_check(hipEventCreate(&b), "hipEventCreate");
            it = events.emplace(name, std::make_pair(a, b)).first;
        }
        hip_check(hipEventRecord(it->second.first, nullptr), "hipEventRecord");
    }
    void stop(const std::string &name) override {
        hip_check(hipEventRecord(events[name].second, nullptr), "hipEventRecord");
    }
    void synchronize() override {
        for (auto &kv : events) {
            float ms = 0;
            hip_check(hipEventSynchronize(kv.second.second), "hipEventSynchronize");
            hip_check(hipEventElapsedTime(&ms, kv.second.first, kv.second.second), "hipEventElapsedTime");
            auto &t = totals()[kv.first];
            t.first += ms;
            t.second += 1;
            (void)hipEventDestroy(kv.second.first);
            (void)hipEventDestroy(kv.second.second);
        }
        events.clear();
    }
    ~event_timer() override {
        for (auto &kv : events) {
            (void)hipEventDestroy(kv.second.first);
            (void)hipEventDestroy(kv.second.second);
        }
    }
};

#define time_this_wf_step wf::raii_timer local_wf_raii_timer(id_str, *pf->timer)

// ---------------------------------------------------------------- raydata

raydata::raydata(int w_, int h_) { resize(w_, h_); }

void raydata::resize(int w_, int h_) {
    w = w_;
    h = h_;
    rays.resize(32ull * n());
    hits.resize(16ull * n());
    occluded.resize(n());
}

hiprt_rays raydata::in() const {
    const float *b = rays.as<const float>();
    size_t m = n();
    return hiprt_rays{b, b + m, b + 2 * m, b + 3 * m, b + 4 * m, b + 5 * m, b + 6 * m, b + 7 * m};
}
hiprt_ray_buffers raydata::out() const {
    float *b = rays.as<float>();
    size_t m = n();
    return hiprt_ray_buffers{b, b + m, b + 2 * m, b + 3 * m, b + 4 * m, b + 5 * m, b + 6 * m, b + 7 * m};
}
hiprt_hits raydata::hit() const {
    float *b = hits.as<float>();
    size_t m = n();
    return hiprt_hits{b, b + m, b + 2 * m, reinterpret_cast<uint32_t *>(b + 3 * m)};
}

// ---------------------------------------------------------------- tracers

void batch_rt::compute_closest_hit() {
    if (!rd) throw std::logic_error("tracer used without ray data");
    if (!pf->ds) throw std::logic_error("no committed scene");
    hiprt_rays r = rd->in();
    hiprt_hits h = rd->hit();
    check(hiprt_trace_closest(pf->ds, &r, rd->n(), &h, tracer, &pf->opts, nullptr), "find closest hit");
}

void batch_rt::compute_any_hit() {
    if (!rd) throw std::logic_error("tracer used without ray data");
    if (!pf->ds) throw std::logic_error("no committed scene");
    hiprt_rays r = rd->in();
    hiprt_hits h = rd->hit();
    check(hiprt_trace_any(pf->ds, &r, rd->n(), rd->occluded.as<uint8_t>(), &h, tracer, &pf->opts, nullptr),
          "find any hit");
}

// ---------------------------------------------------------------- steps

struct rd_step : public wf::step {
    raydata *rd = nullptr;
    void use(raydata *r) { rd = r; }
};

struct initialize_framebuffer : public wf::step {
    static constexpr char id[] = "initialize framebuffer";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        pf->framebuffer.resize(16ull * rc->w * rc->h);
        hip_check(hipMemsetAsync(pf->framebuffer.p, 0, pf->framebuffer.bytes, nullptr), "hipMemsetAsync");
    }
};

struct sample_camera_rays : public rd_step {
    static constexpr char id[] = "sample camera rays";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        hiprt_ray_buffers o = rd->out();
        if (pf->rt && pf->rt->use_incoherence) {
            // rt/cuda/rni.cpp:36-89: two spheres along the scene's largest
            // extent, a quarter in from either end; the generator restarts
            // from one seed (1234) for every sample, as the reference's does
            float s1[3], s2[3], r_max;
            incoherence_spheres(s1, s2, r_max);
            check(hiprt_incoherent_rays(s1, s2, pf->rt->incoherence_r1, pf->rt->incoherence_r2, r_max, 1234, &o,
                                        rd->n(), nullptr),
                  id);
            return;
        }
        float cam[12];
        rc->camera(cam);
        check(hiprt_camera_rays(cam, &o, nullptr), id);
    }
    // Sphere centres and r_max from the extent of the scene's vertices.
    static void incoherence_spheres(float *s1, float *s2, float &r_max) {
        uint32_t nv = 0;
        check(hiprt_scene_info(rc->scene.geometry, &nv, nullptr), id);
        std::vector<float> pos(3ull * nv);
        check(hiprt_scene_copy(rc->scene.geometry, pos.data(), nullptr, nullptr), id);
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) {
                if (pos[3 * i + k] < lo[k]) lo[k] = pos[3 * i + k];
                if (pos[3 * i + k] > hi[k]) hi[k] = pos[3 * i + k];
            }
        float size[3];
        for (int k = 0; k < 3; ++k) size[k] = std::fabs(hi[k] - lo[k]);
        const int a = (size[0] >= size[1] && size[0] >= size[2]) ? 0 : (size[1] >= size[2] ? 1 : 2);
        for (int k = 0; k < 3; ++k) s1[k] = s2[k] = lo[k] + size[k] / 2;
        s1[a] = lo[a] + size[a] / 4;
        s2[a] = hi[a] - size[a] / 4;
        r_max = size[a] / 4;
    }
};

struct find_closest_hits : public wf::find_closest_hits {
    std::string id_str = id;
    raydata *rd = nullptr;
    find_closest_hits() : wf::find_closest_hits(nullptr) {}
    void use(wf::raydata *r) override { rd = dynamic_cast<raydata *>(r); }
    void run() override {
        time_this_wf_step;
        pf->rt->use(rd);
        pf->rt->compute_closest_hit();
    }
};

struct find_any_hits : public wf::find_any_hits {
    std::string id_str = id;
    raydata *rd = nullptr;
    find_any_hits() : wf::find_any_hits(nullptr) {}
    void use(wf::raydata *r) override { rd = dynamic_cast<raydata *>(r); }
    void run() override {
        time_this_wf_step;
        pf->rt->use(rd);
        pf->rt->compute_any_hit();
    }
};

struct add_hitpoint_albedo : public rd_step {
    static constexpr char id[] = "add hitpoint albedo";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        hiprt_hits h = rd->hit();
        check(hiprt_add_hitpoint_albedo(pf->ds, &h, rd->n(), pf->framebuffer.as<float>(), nullptr), id);
    }
};

// Per-sample data of the direct-lighting algorithm (per_sample_data<float> /
// <vec3> of gi/direct.cpp:478-481): the sample pdf and the light colour.
struct direct_data {
    device_buffer pdf, light_col;
    void resize(uint32_t n) {
        pdf.resize(4ull * n);
        light_col.resize(12ull * n);
    }
};

// compute_light_distribution (rt/cpu/preprocessing.cpp:9-20): the light list
// of the committed scene -- its emissive triangles, then the point lights of
// the script -- with the power distribution.
struct compute_light_distribution : public wf::step {
    static constexpr char id[] = "compute light distribution";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        std::vector<float> pts;
        for (const pointlight &l : rc->scene.lights) pts.insert(pts.end(), {l.pos[0], l.pos[1], l.pos[2], l.col[0], l.col[1], l.col[2]});
        uint32_t n = 0;
        check(hiprt_scene_set_lights(pf->ds, pts.data(), (uint32_t)(pts.size() / 6), &n), id);
        if (n == 0) std::cerr << "WARNING: There is neither emissive geometry nor a light" << std::endl;
    }
};

// sample_uniform_dir, sample_cos_weighted_dir, sample_light_dir
// (rt/cpu/bounce.cpp:11-106): one shadow/direction ray per camera hit, its
// pdf (and the light colour); camera hits on emitters add their emission to
// the framebuffer.  Sample k of the step draws from seed k (1, 2, ...).
template <int MODE, const char *ID> struct direct_sample_step : public wf::step {
    static constexpr const char *id = ID;
    std::string id_str = id;
    raydata *cam = nullptr, *shadow = nullptr;
    direct_data *dd = nullptr;
    uint64_t seed = 1;
    void run() override {
        time_this_wf_step;
        dd->resize(cam->n());
        hiprt_rays r = cam->in();
        hiprt_hits h = cam->hit();
        hiprt_ray_buffers o = shadow->out();
        check(hiprt_sample_direct(pf->ds, MODE, &r, &h, cam->n(), seed++, &o, dd->pdf.as<float>(),
                                  dd->light_col.as<float>(), pf->framebuffer.as<float>(), nullptr),
              id);
    }
};
constexpr char kSampleUniform[] = "sample uniform dir";
constexpr char kSampleCosine[] = "sample cos weighted dir";
constexpr char kSampleLight[] = "sample light dir";
using sample_uniform_dir = direct_sample_step<HIPRT_DIRECT_UNIFORM, kSampleUniform>;
using sample_cos_weighted_dir = direct_sample_step<HIPRT_DIRECT_COSINE, kSampleCosine>;
using sample_light_dir = direct_sample_step<HIPRT_DIRECT_LIGHT, kSampleLight>;

struct integrate_dir_sample : public wf::step {
    static constexpr char id[] = "integrate dir sample";
    std::string id_str = id;
    raydata *cam = nullptr, *shadow = nullptr;
    direct_data *dd = nullptr;
    void run() override {
        time_this_wf_step;
        hiprt_rays cr = cam->in(), sr = shadow->in();
        hiprt_hits ch = cam->hit(), sh = shadow->hit();
        check(hiprt_integrate_dir_sample(pf->ds, &cr, &ch, &sr, &sh, dd->pdf.as<float>(), cam->n(),
                                         pf->framebuffer.as<float>(), nullptr),
              id);
    }
};

struct integrate_light_sample : public wf::step {
    static constexpr char id[] = "integrate light sample";
    std::string id_str = id;
    raydata *cam = nullptr, *shadow = nullptr;
    direct_data *dd = nullptr;
    void run() override {
        time_this_wf_step;
        hiprt_rays cr = cam->in(), sr = shadow->in();
        hiprt_hits ch = cam->hit();
        check(hiprt_integrate_light_sample(pf->ds, &cr, &ch, &sr, shadow->occluded.as<uint8_t>(),
                                           dd->pdf.as<float>(), dd->light_col.as<float>(), cam->n(),
                                           pf->framebuffer.as<float>(), nullptr),
              id);
    }
};

struct download_framebuffer : public wf::step {
    static constexpr char id[] = "download framebuffer";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        size_t n = (size_t)rc->w * rc->h;
        rc->framebuffer.resize(4 * n);
        hip_check(hipMemcpy(rc->framebuffer.data(), pf->framebuffer.p, 16 * n, hipMemcpyDeviceToHost), "hipMemcpy");
        for (size_t i = 0; i < n; ++i) {
            float wgt = rc->framebuffer[4 * i + 3];
            if (wgt > 0)
                for (int k = 0; k < 4; ++k) rc->framebuffer[4 * i + k] /= wgt;
        }
    }
};

struct build_accel_struct : public wf::build_accel_struct {
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        pf->build();
    }
};

// The reference's example scene steps (rt/cuda/preprocessing.cpp:41-72,
// preprocessing.h:29-52).  "rotate scene" turns the vertex positions by
// Rz(pi/4) Ry(pi/4) Rx(pi/4) about the origin, for the steps that follow it
// at commit (normally "build accel struct"); normals stay as they are, as
// the reference's kernel is given no normal buffer.  "rotate scene keeping
// original data" keeps the positions it replaces, and "drop scene view" puts
// the last kept ones back for the next commit (a device scene built in
// between keeps the turned ones).
struct rotate_scene : public wf::step {
    static constexpr char id[] = "rotate scene";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        pf->rotate_scene(false);
    }
};
struct rotate_scene_keep_org : public wf::step {
    static constexpr char id[] = "rotate scene keeping original data";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        pf->rotate_scene(true);
    }
};
struct drop_scene_view : public wf::step {
    static constexpr char id[] = "drop scene view";
    std::string id_str = id;
    void run() override {
        time_this_wf_step;
        pf->drop_scene_view();
    }
};

// The rotation of the example steps, rot = m_z * m_y * m_x with
// m_a = rotate(pi/4, a) (Rodrigues' formula about the unit axis a), in float
// and in the reference's matrix library's summation order; R[column][row].
void rotation_xyz_quarter(float R[3][3]) {
    const float angle = 3.14159265358979323846f / 4;
    const float c = std::cos(angle), s = std::sin(angle);
    float M[3][3][3];   // M[axis][column][row]
    for (int a = 0; a < 3; ++a) {
        float ax[3] = {0, 0, 0};
        ax[a] = 1;
        const float t[3] = {(1.0f - c) * ax[0], (1.0f - c) * ax[1], (1.0f - c) * ax[2]};
        M[a][0][0] = c + t[0] * ax[0];
        M[a][0][1] = t[0] * ax[1] + s * ax[2];
        M[a][0][2] = t[0] * ax[2] - s * ax[1];
        M[a][1][0] = t[1] * ax[0] - s * ax[2];
        M[a][1][1] = c + t[1] * ax[1];
        M[a][1][2] = t[1] * ax[2] + s * ax[0];
        M[a][2][0] = t[2] * ax[0] + s * ax[1];
        M[a][2][1] = t[2] * ax[1] - s * ax[0];
        M[a][2][2] = c + t[2] * ax[2];
    }
    auto mul = [](const float A[3][3], const float B[3][3], float out[3][3]) {
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r)
                out[col][r] = (A[0][r] * B[col][0] + A[1][r] * B[col][1]) + A[2][r] * B[col][2];
    };
    float zy[3][3];
    mul(M[2], M[1], zy);
    mul(zy, M[0], R);
}

void platform::rotate_scene(bool keep) {
    if (!sd || !sd->geometry) throw std::logic_error("There is no scene data to work with");
    uint32_t nv = 0, nt = 0;
    check(hiprt_scene_info(sd->geometry, &nv, &nt), "rotate scene");
    std::vector<float> pos(3ull * nv);
    check(hiprt_scene_copy(sd->geometry, pos.data(), nullptr, nullptr), "rotate scene");
    if (keep) views.push_back(pos);
    float R[3][3];
    rotation_xyz_quarter(R);
    for (size_t i = 0; i < nv; ++i) {
        const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        for (int r = 0; r < 3; ++r) pos[3 * i + r] = (x * R[0][r] + y * R[1][r]) + z * R[2][r];
    }
    check(hiprt_scene_set_positions(sd->geometry, pos.data(), nv), "rotate scene");
}

void platform::drop_scene_view() {
    if (views.empty()) throw std::logic_error("drop scene view: there is no scene view to drop");
    if (!sd || !sd->geometry) throw std::logic_error("There is no scene data to work with");
    uint32_t nv = 0;
    check(hiprt_scene_info(sd->geometry, &nv, nullptr), "drop scene view");
    if (3ull * nv != views.back().size())
        throw std::logic_error("drop scene view: the scene's vertex count changed since its view was made");
    check(hiprt_scene_set_positions(sd->geometry, views.back().data(), nv), "drop scene view");
    views.pop_back();
}

// ---------------------------------------------------------------- platform

platform::platform(const std::vector<std::string> &args) : wf::platform("hip") {
    if (pf) throw std::logic_error("The " + name + " platform is already set up");
    for (auto &a : args) std::cerr << "Platform hip does not support the argument " << a << std::endl;
    check(hiprt_create(0, &ctx), "platform hip");
    pf = this;
    // every tracer and its alpha-aware "-alpha" twin (rt/cuda/platform.cpp:107-122)
    for (int a : {0, (int)HIPRT_TRACER_ALPHA})
        for (int t = 0; t < HIPRT_TRACER_COUNT; ++t) {
            int id = t | a;
            std::string n = hiprt_tracer_name(id);
            tracers[n] = [id]() -> wf::batch_ray_tracer * { return new batch_rt(id); };
        }
    // "default" and "find closest hits" link to while-while as in
    // rt/cuda/platform.cpp:55-56; the reference's other CUDA loop structures
    // of the same near-first traversal (rt/cuda/platform.cpp:107-122) select
    // the wave64 kernel that runs it, so scripts naming them run unchanged
    // (hiprt_tracer_id agrees)
    link_tracer(hiprt_tracer_name(hiprt_tracer_id("find closest hits")), "find closest hits");
    for (const char *alias : {"default", "simple", "if-if", "persistent-if-if", "speculative-while-while"})
        for (const char *suffix : {"", "-alpha"}) {
            std::string n = std::string(alias) + suffix;
            link_tracer(hiprt_tracer_name(hiprt_tracer_id(n.c_str())), n);
        }
    register_wf_step_by_id(, initialize_framebuffer);
    register_wf_step_by_id(, sample_camera_rays);
    register_wf_step_by_id(, hip::find_closest_hits);
    register_wf_step_by_id(, hip::find_any_hits);
    register_wf_step_by_id(, add_hitpoint_albedo);
    register_wf_step_by_id(, compute_light_distribution);
    register_wf_step_by_id(, sample_uniform_dir);
    register_wf_step_by_id(, sample_cos_weighted_dir);
    register_wf_step_by_id(, sample_light_dir);
    register_wf_step_by_id(, integrate_dir_sample);
    register_wf_step_by_id(, integrate_light_sample);
    register_wf_step_by_id(, download_framebuffer);
    register_wf_step_by_id(, hip::build_accel_struct);
    register_wf_step_by_id(, hip::rotate_scene);
    register_wf_step_by_id(, hip::rotate_scene_keep_org);
    register_wf_step_by_id(, hip::drop_scene_view);
    timer = new event_timer;
}

platform::~platform() {
    (void)hipDeviceSynchronize();
    if (ds) hiprt_scene_destroy(ds);
    if (bvh) hiprt_bvh_free(bvh);
    framebuffer.resize(0);
    if (ctx) hiprt_destroy(ctx);
    pf = nullptr;
}

raydata *platform::allocate_raydata() { return new raydata(rc->w, rc->h); }

void platform::commit_scene(wf::scene *s) {
    sd = s;
    if (!rt) rt = dynamic_cast<batch_rt *>(select("default"));
    for (auto *st : scene_steps) st->run();
}

void platform::build() {
    if (!sd || !sd->geometry) throw std::logic_error("There is no scene data to work with");
    if (ds) hiprt_scene_destroy(ds);
    if (bvh) hiprt_bvh_free(bvh);
    ds = nullptr;
    bvh = nullptr;
    check(hiprt_bvh_build(sd->geometry, &bvh_params, &bvh), "build accel struct");
    check(hiprt_scene_commit(ctx, sd->geometry, bvh, HIPRT_BVH2, tri_format, &ds), "build accel struct");
}

#define error(x)                                                                  \
    {                                                                             \
        std::cerr << "command (" << command << "): " << x << std::endl;            \
        return true;                                                              \
    }
#define check_in_complete(x)                                                      \
    {                                                                             \
        if (in.bad() || in.fail() || !in.eof()) error(x);                         \
    }

// "raytracer <variant>" (rt/cuda/platform.cpp:102-131) and the BVH settings
// of batch_rt::interprete (rt/cuda/base.cpp:119-155), which here really are
// routed to the platform.
bool platform::interprete(const std::string &command, std::istringstream &in) {
    if (command == "raytracer") {
        std::string variant;
        in >> variant;
        check_in_complete("Syntax error, requires a hip ray tracer variant name");
        if (tracers.count(variant) == 0) error("There is no such hip ray tracer variant");
        rt = dynamic_cast<batch_rt *>(select(variant));
        return true;
    }
    if (command == "incoherence") {
        // batch_rt::interprete (rt/cuda/base.cpp:120-133): a setting of the
        // current tracer; "0 0" leaves an earlier one in place
        float r1 = 0, r2 = 0;
        in >> r1;
        in >> r2;
        check_in_complete("Syntax error, \"incoherence\" requires exactly two positive float values");
        if (r1 < 0 || r2 < 0) error("Parameter error, \"incoherence\" requires exactly two positive float values");
        if (!rt) rt = dynamic_cast<batch_rt *>(select("default"));
        if (r1 != 0 || r2 != 0) {
            rt->use_incoherence = true;
            rt->incoherence_r1 = r1;
            rt->incoherence_r2 = r2;
        }
        return true;
    }
    if (command == "bvh") {
        std::string sub;
        in >> sub;
        if (sub == "type") {
            std::string t;
            in >> t;
            check_in_complete("Syntax error, \"bvh type\" requires exactly one string value");
            if (t == "om") bvh_params.split = 0;
            else if (t == "sm") bvh_params.split = 1;
            else if (t == "sah") bvh_params.split = 2;
            else error("Parameter error, \"bvh type\" must be one of \"sm\", \"om\", \"sah\"");
            return true;
        }
        if (sub == "max_tris" || sub == "planes") {
            int v;
            in >> v;
            check_in_complete("Syntax error, \"bvh " << sub << "\" requires exactly one positive integer value");
            if (v < 1) error("Parameter error, value must be positive");
            (sub == "max_tris" ? bvh_params.max_tris : bvh_params.planes) = v;
            return true;
        }
        if (sub == "esc") {
            std::string v;
            in >> v;
            check_in_complete("Syntax error, \"bvh esc\" requires on or off");
            if (v != "on" && v != "off") error("Syntax error, \"bvh esc\" requires on or off");
            bvh_params.esc = v == "on";
            return true;
        }
        error("Unknown bvh subcommand " << sub);
    }
    if (command == "tracer-option") {
        // No counterpart in the reference (its CUDA tracers are compiled with
        // fixed settings): the persistent tracers' tunables.
        std::string name;
        int v = 0;
        in >> name >> v;
        check_in_complete("Syntax error, \"tracer-option\" requires a name and an integer value");
        if (name == "refill") {
            if (v < 1 || v > 64) error("Parameter error, refill threshold must be 1..64");
            opts.refill_threshold = v;
        } else if (name == "blocks") {
            if (v < 0) error("Parameter error, blocks must be >= 0 (0 = occupancy default)");
            opts.blocks = v;
        } else if (name == "stack") {
            if (v < 0 || v > 4) error("Parameter error, stack configuration must be 0..4");
            opts.stack_config = v;
        } else if (name == "chunks") {
            if (v != 1 && v != 8) error("Parameter error, chunks must be 1 or 8");
            opts.work_chunks = v;
        } else if (name == "ray-order") {
            if (v != 0 && v != 1) error("Parameter error, ray-order must be 0 (sorted) or 1 (shading order)");
            ray_order = v;
        } else if (name == "samples") {
            if (v < 1 || v > 64) error("Parameter error, samples must be 1..64");
            samples = v;
        } else if (name == "park") {
            if (v < 1 || v > 4) error("Parameter error, park must be 1..4");
            opts.park_leaves = v;
        } else if (name == "regen") {
            if (v < 0 || v > 64) error("Parameter error, regen must be 0..64 (path slots per pixel, 0 = off)");
            regen_pool = v;
        } else {
            error("Unknown tracer option " << name << " (refill, blocks, stack, chunks, ray-order, samples, park, regen)");
        }
        return true;
    }
    if (command == "triangles") {
        std::string v;
        in >> v;
        check_in_complete("Syntax error, \"triangles\" requires exact or woop");
        if (v == "exact") tri_format = HIPRT_TRI_EXACT;
        else if (v == "woop") tri_format = HIPRT_TRI_WOOP;
        else error("Syntax error, \"triangles\" requires exact or woop");
        return true;
    }
    return false;
}

#undef error
#undef check_in_complete

// ---------------------------------------------------------------- algorithms

namespace {

// primary-wf: camera rays, closest hit, hit-point albedo (gi/primary-hit.cpp:85-117).
class primary_wf : public simple_algorithm {
    std::unique_ptr<raydata> rd;

public:
    primary_wf() {
        rd.reset(pf->allocate_raydata());
        auto *cam = pf->step<sample_camera_rays>();
        auto *hit = pf->step<hip::find_closest_hits>();
        auto *alb = pf->step<add_hitpoint_albedo>();
        cam->use(rd.get());
        hit->use(rd.get());
        alb->use(rd.get());
        frame_preparation_steps.push_back(pf->step<initialize_framebuffer>());
        sampling_steps = {cam, hit, alb};
        frame_finalization_steps.push_back(pf->step<download_framebuffer>());
    }
};

// direct-wf (wf::direct_light, gi/direct.cpp:470-585): camera rays and their
// closest hits, then by the "is" mode
//   uniform (the default) / cosine: sample a direction, trace its closest hit
//     ("secondary hits"), integrate dir sample;
//   light: compute light distribution (a data-reset step), sample light dir,
//     any hit ("find occluders"), integrate light sample.
class direct_wf : public simple_algorithm {
    std::unique_ptr<raydata> cam_rd, shadow_rd;
    direct_data dd;
    int mode = HIPRT_DIRECT_UNIFORM;

    template <class S> void wire_sample(S *s) {
        s->cam = cam_rd.get();
        s->shadow = shadow_rd.get();
        s->dd = &dd;
    }
    void regenerate_steps() {
        data_reset_steps.clear();
        frame_preparation_steps = {pf->step<initialize_framebuffer>()};
        frame_finalization_steps = {pf->step<download_framebuffer>()};
        auto *cam = pf->step<sample_camera_rays>("primary hits");
        auto *hit = pf->step<hip::find_closest_hits>();
        cam->use(cam_rd.get());
        hit->use(cam_rd.get());
        if (mode == HIPRT_DIRECT_LIGHT) {
            data_reset_steps.push_back(pf->step<compute_light_distribution>());
            auto *sample = pf->step<sample_light_dir>();
            auto *occ = pf->step<hip::find_any_hits>("find occluders");
            auto *integ = pf->step<integrate_light_sample>();
            wire_sample(sample);
            occ->use(shadow_rd.get());
            wire_sample(integ);
            sampling_steps = {cam, hit, sample, occ, integ};
        } else {
            wf::step *sample;
            if (mode == HIPRT_DIRECT_UNIFORM) {
                auto *u = pf->step<sample_uniform_dir>();
                wire_sample(u);
                sample = u;
            } else {
                auto *c = pf->step<sample_cos_weighted_dir>();
                wire_sample(c);
                sample = c;
            }
            auto *trace = pf->step<hip::find_closest_hits>("secondary hits");
            auto *integ = pf->step<integrate_dir_sample>();
            trace->use(shadow_rd.get());
            wire_sample(integ);
            sampling_steps = {cam, hit, sample, trace, integ};
        }
    }

public:
    direct_wf() {
        cam_rd.reset(pf->allocate_raydata());
        shadow_rd.reset(pf->allocate_raydata());
        regenerate_steps();
    }
    // "is uniform|cosine|light" (gi/direct.cpp:570-582)
    bool interprete(const std::string &command, std::istringstream &in) override {
        if (command != "is") return false;
        std::string value;
        in >> value;
        if (value == "uniform") mode = HIPRT_DIRECT_UNIFORM;
        else if (value == "cosine") mode = HIPRT_DIRECT_COSINE;
        else if (value == "light") mode = HIPRT_DIRECT_LIGHT;
        else if (value == "brdf") throw std::runtime_error("unsupported importance sampling method for wf/direct");
        else std::cerr << "unknown sampling mode in interprete: " << value << std::endl;
        regenerate_steps();
        return true;
    }
};

// pt-wf: the batched wavefront path tracer of the library, one frame per
// sample, the accumulator averaged into the framebuffer at the end.  With
// "tracer-option samples k" each render call traces k samples per pixel
// together (k must divide sppx), so sppx / k calls render the same sppx
// frames.  "tracer-option regen k" renders them with path regeneration over
// k path slots per pixel instead of all k samples in flight.
class pt_wf : public simple_algorithm {
    hiprt_pt pt = nullptr;
    int frames = 0;      // render calls so far
    int samples = 1;     // samples per render call
    int max_depth = 4;   // last trace level (levels 0..max_depth)
    int rr_start = 3;    // first level whose ray is subject to Russian roulette

public:
    // "path len|bounce|rr-start" of simple_pt (gi/pt.cpp:121-154), with its
    // messages.  The reference's loop traces rays i = 0 .. len-1 and, shading
    // the hit of ray i > rr-start, plays roulette on ray i + 1: here "len n"
    // traces levels 0 .. n-1 and "rr-start n" subjects the rays of levels
    // n + 2 and later.  Bounces sample the Lambertian BRDF, whose importance
    // sampling is the cosine distribution: "bounce brdf" and "bounce cosine"
    // both select it.
    bool interprete(const std::string &command, std::istringstream &in) override {
        if (command != "path") return false;
        std::string sub, val;
        in >> sub;
        if (sub == "len") {
            int i = 0;
            in >> i;
            if (i <= 0) std::cerr << "error in path len: expected a positive integer, got " << i << std::endl;
            else if (i > 65) std::cerr << "error in path len: pt-wf traces at most 65 levels, got " << i << std::endl;
            else max_depth = i - 1;
        } else if (sub == "bounce") {
            in >> val;
            if (val == "uniform") std::cerr << "error: pt-wf samples bounces by the BRDF (cosine) only" << std::endl;
            else if (val != "cosine" && val != "brdf")
                std::cerr << "error: invalid kind of path bounce: '" << val << "'" << std::endl;
        } else if (sub == "rr-start") {
            int i = 0;
            in >> i;
            if (i <= 0) std::cerr << "error in russian roulette offset: expected a positive integer > 0, got " << i << std::endl;
            else rr_start = i + 2;
        } else {
            std::cerr << "unknown subcommand to path: '" << sub << "'" << std::endl;
        }
        return true;
    }
    ~pt_wf() override {
        if (pt) hiprt_pt_destroy(pt);
    }
    void prepare_frame() override {
        simple_algorithm::prepare_frame();
        if (pt) hiprt_pt_destroy(pt);
        pt = nullptr;
        float cam[12];
        rc->camera(cam);
        int tracer = pf->rt ? pf->rt->tracer : -1;
        // batched needs a persistent tracer: the plain one becomes the default (keeping "-alpha")
        if (tracer >= 0 && (tracer & ~HIPRT_TRACER_ALPHA) == HIPRT_TRACER_WHILE_WHILE)
            tracer = HIPRT_TRACER_DYNAMIC_WHILE_WHILE | (tracer & HIPRT_TRACER_ALPHA);
        const hiprt_tracer_opts &o = pf->opts;
        samples = pf->samples;
        if (rc->sppx % samples != 0)
            throw std::runtime_error("pt-wf: sppx " + std::to_string(rc->sppx) +
                                     " is not a multiple of tracer-option samples " + std::to_string(samples));
        hiprt_pt_params p{max_depth, rr_start, 0.8f, 1, tracer, 1, o.stack_config, o.work_chunks, o.refill_threshold,
                          o.blocks, pf->ray_order, samples, 0, o.park_leaves, 0, pf->regen_pool};
        check(hiprt_pt_create(pf->ds, cam, &p, &pt), "pt-wf");
        frames = 0;
    }
    bool compute_sample(int sppx, wf::timer *t) override {
        if (current_sample_index >= sppx) return false;
        current_sample_index += samples;
        t->start("path trace frame");
        check(hiprt_pt_render(pt, (uint32_t)frames++, nullptr), "pt-wf");
        t->stop("path trace frame");
        t->synchronize();
        return current_sample_index < sppx;
    }
    void finalize_frame(wf::timer *t) override {
        size_t n = (size_t)rc->w * rc->h;
        std::vector<float> rgb(3 * n);
        check(hiprt_pt_image(pt, rgb.data()), "pt-wf");
        rc->framebuffer.assign(4 * n, 1.0f);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                rc->framebuffer[4 * i + k] = rgb[3 * i + k] / (float)(frames ? frames * samples : 1);
        if (t) t->synchronize();
    }
};

} // namespace

void print_stats(std::ostream &out) {
    for (auto &kv : event_timer::totals())
        out << kv.first << ": " << kv.second.first << " ms in " << kv.second.second << " calls ("
            << kv.second.first / (kv.second.second ? kv.second.second : 1) << " ms avg)" << std::endl;
}

void clear_stats() { event_timer::totals().clear(); }

std::unique_ptr<simple_algorithm> make_algorithm(const std::string &name) {
    if (!pf) throw std::logic_error("wavefront algorithms need the hip platform");
    if (name == "primary-wf") return std::make_unique<primary_wf>();
    if (name == "direct-wf") return std::make_unique<direct_wf>();
    if (name == "pt-wf") return std::make_unique<pt_wf>();
    return nullptr;
}

} // namespace hip
} // namespace wf
