// The wavefront platform interface of RTGI, restated for the HIP platform
// shim: step and tracer registries with name caching (select, link_tracer,
// step), the batch tracer interface (compute_closest_hit, compute_any_hit,
// use), the trace steps, the setup-step list and the simple step-list
// algorithm.  Names, argument meaning and error behaviour follow
// libgi/wavefront-rt.h:22-186 and libgi/wavefront-rt.cpp:4-45, so a platform
// written against the reference's interface reads the same here.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace wf {

// Asynchronous step timer: start/stop may only enqueue; synchronize() must
// resolve the measurements (wavefront-rt.h:22-24).
struct timer {
    virtual ~timer() {}
    virtual void start(const std::string &name) = 0;
    virtual void stop(const std::string &name) = 0;
    virtual void synchronize() = 0;
};

struct raii_timer {
    const std::string &name;
    timer &t;
    raii_timer(const std::string &n, timer &tm) : name(n), t(tm) { t.start(name); }
    ~raii_timer() { t.stop(name); }
};

struct step {
    virtual ~step() {}
    virtual void run() = 0;
};

struct raydata {
    virtual ~raydata() {}
};

// Any ray tracer may take commands (libgi/rt.h:105-111).
struct ray_tracer {
    virtual ~ray_tracer() {}
    virtual bool interprete(const std::string &command, std::istringstream &in) { return false; }
};

struct batch_ray_tracer : public ray_tracer {
    virtual void compute_closest_hit() = 0;
    virtual void compute_any_hit() = 0;
    virtual void use(raydata *) = 0;
};

struct scene;   // platform-independent scene description (driver side)

#define register_batch_rt(N, C, X) tracers[N] = [C]() -> wf::batch_ray_tracer * { return new X; }
#define register_wf_step(N, C, X) steps[N] = [C]() -> wf::step * { return new X; }
#define register_wf_step_by_id(C, X) steps[X::id] = [C]() -> wf::step * { return new X; }

class platform {
protected:
    std::map<std::string, std::function<batch_ray_tracer *()>> tracers;
    std::map<std::string, std::function<wf::step *()>> steps;
    std::set<std::string> tracer_links;
    std::map<std::string, batch_ray_tracer *> generated_tracers;
    std::map<std::string, wf::step *> generated_steps;
    std::vector<wf::step *> scene_steps;

    void link_tracer(const std::string &existing, const std::string &linkname) {
        tracer_links.insert(linkname);
        tracers[linkname] = tracers[existing];
    }

public:
    std::string name;
    batch_ray_tracer *selected_tracer = nullptr;
    wf::timer *timer = nullptr;

    explicit platform(const std::string &n) : name(n) {}
    virtual ~platform() {
        for (auto &kv : generated_steps) delete kv.second;
        std::set<batch_ray_tracer *> seen;
        for (auto &kv : generated_tracers)
            if (seen.insert(kv.second).second) delete kv.second;
        delete timer;
    }

    // Cached per name; throws std::logic_error for an unknown tracer
    // (wavefront-rt.cpp:19-30).
    batch_ray_tracer *select(const std::string &tname) {
        auto it = generated_tracers.find(tname);
        if (it != generated_tracers.end()) {
            selected_tracer = it->second;
        } else {
            if (tracers.count(tname) == 0)
                throw std::logic_error("No tracer " + tname + " for platform " + this->name);
            selected_tracer = tracers[tname]();
            generated_tracers[tname] = selected_tracer;
        }
        return selected_tracer;
    }

    // Cached per instance name; nullptr for an unknown step id.
    wf::step *step(const std::string &id, std::string sname = "") {
        if (sname.empty()) sname = id;
        auto it = generated_steps.find(sname);
        if (it != generated_steps.end()) return it->second;
        if (steps.count(id) == 0) return nullptr;
        wf::step *s = steps[id]();
        generated_steps[sname] = s;
        return s;
    }
    template <typename T> T *step(const std::string &sname = T::id) {
        return dynamic_cast<T *>(step(T::id, sname));
    }
    std::vector<std::string> tracer_names() const {
        std::vector<std::string> v;
        for (auto &kv : tracers) v.push_back(kv.first);
        return v;
    }

    virtual raydata *allocate_raydata() = 0;
    virtual void commit_scene(scene *s) = 0;
    virtual bool interprete(const std::string &command, std::istringstream &in) { return false; }
    void append_setup_step(wf::step *s) { scene_steps.push_back(s); }
};

// The trace steps (wavefront-rt.h:152-178).
class find_closest_hits : public step {
protected:
    batch_ray_tracer *rt;

public:
    static constexpr char id[] = "find closest hit";
    explicit find_closest_hits(batch_ray_tracer *r) : rt(r) {}
    virtual void use(raydata *) = 0;
};
class find_any_hits : public step {
protected:
    batch_ray_tracer *rt;

public:
    static constexpr char id[] = "find any hit";
    explicit find_any_hits(batch_ray_tracer *r) : rt(r) {}
    virtual void use(raydata *) = 0;
};
struct build_accel_struct : public step {
    static constexpr char id[] = "build accel struct";
};

// Step-list algorithm (wavefront-rt.h:112-142): per sample, every sampling
// step in order, then the platform timer is synchronized.
class simple_algorithm {
protected:
    std::vector<step *> data_reset_steps, frame_preparation_steps, sampling_steps, frame_finalization_steps;
    int current_sample_index = 0;

public:
    virtual ~simple_algorithm() {}
    virtual void prepare_data() {
        for (auto *s : data_reset_steps) s->run();
    }
    virtual void prepare_frame() {
        current_sample_index = 0;
        for (auto *s : frame_preparation_steps) s->run();
    }
    virtual bool compute_sample(int sppx, wf::timer *t) {
        if (current_sample_index >= sppx) return false;
        ++current_sample_index;
        for (auto *s : sampling_steps) s->run();
        if (t) t->synchronize();
        return current_sample_index < sppx;
    }
    virtual void finalize_frame(wf::timer *t) {
        for (auto *s : frame_finalization_steps) s->run();
        if (t) t->synchronize();
    }
    // Algorithm-specific commands (algorithm::interprete, libgi/algorithm.h).
    virtual bool interprete(const std::string &command, std::istringstream &in) { return false; }
};

} // namespace wf
