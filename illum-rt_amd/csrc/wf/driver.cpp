// rtgi-hip: RTGI's script front end for the commands on the batch-tracing
// path (driver/interaction.cpp:247-856), driving the hip platform.  A script
// such as
//
//     platform hip
//     raytracer dynamic-while-while
//     load scene.obj            (or: generate 3)
//     add-scene-step build accel struct
//     commit
//     algo primary-wf
//     run
//
// runs as it would on RTGI's cuda platform.  Errors print
// "command <n> (<name>): <message>" (n counts the script's lines; the
// platform's own messages have no number, as the reference's platform prints
// them) and the script continues, as in the
// reference; "run" checks the same preconditions.
//
// Commands: at, look, up, camup, fovy, resolution, sppx, load (appends a
// model under the model matrix), modelmatrix, modeltrafo, modelpath,
// generate, pointlight, material, default-brdf, platform, raytracer,
// tracer-option, bvh, triangles, add-scene-step, commit, algo, run, outfile,
// stats, echo, disable, with / }, quit, and the algorithm's own (direct-wf:
// is uniform|cosine|light; pt-wf: path len|rr-start|bounce).  skylight,
// preview-offset, omp and denoise are checked as the reference checks them
// and change nothing here: the reference's wavefront steps do not sample a
// sky either, there is no preview window, no host threading and no
// OpenImageDenoise.
// Materials are Lambertian: "material roughness|ior|brdf" and "default-brdf"
// are accepted and kept for "material show" but do not change the shading.
// "with gl|cuda|optix {" blocks are skipped (those platforms are not built
// here); "with hip {" blocks run unless "disable hip" came first.  Images are
// written as PFM (float) or, for any other extension, 8-bit PPM with a 2.2
// gamma.
#include "hip-platform.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>

using namespace std;

namespace {

struct update_counters {
    int cmdid = 0;
    int scene_touched_at = 0, tracer_touched_at = 0, accel_touched_at = 0;
    bool valid_platform = false;
};

update_counters uc;
unique_ptr<wf::simple_algorithm> algo;
string algo_name;
string outfile = "out.pfm";
bool skip_block = false;                 // inside a "with <missing> {" block
map<string, bool> disabled_features;
int selected_mat = -1;                   // "material select", -1: none
// parameters of the reference's other BRDFs, kept per material for "material show"
struct brdf_params {
    float roughness = 0.1f, ior = 1.3f;
    string brdf;
};
map<int, brdf_params> mat_params;
string default_brdf = "lambert";
bool brdfs_applied = false;   // a model was loaded: scene::add applied the default BRDF to its materials
// camera/scene up state of interaction.cpp:229-233
bool cam_has_up = false, scene_up_set = false;
float scene_up[3] = {0, 1, 0};
// "skylight": the reference's wavefront steps take triangle lights only
// (rt/cuda/preprocessing.cpp:20-28), so the sky is checked and kept, not sampled
bool have_sky = false;
float sky_scale = 1.0f;

// The BRDF names of new_brdf (libgi/material.cpp:326-370).
bool known_brdf(const string &n) {
    static const char *names[] = {"lambert", "mirror", "specular", "dielectric-specular", "thin-dielectric-specular",
                                  "phong", "layered-phong", "gtr2", "layered-gtr2"};
    for (const char *k : names)
        if (n == k) return true;
    return false;
}
// model matrix for "load" (column-major, m[4 * column + row]) and the model
// search paths ("" first: the working directory), driver/interaction.cpp:203
// and libgi/scene.h:164-170
float modelmatrix[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
vector<string> modelpaths{""};

// Column arithmetic of the reference's matrix library for "modeltrafo":
// translate puts m * (v, 1) into the last column, scale scales the first
// three columns, rotate multiplies by the rotation of `angle` about the unit
// axis (Rodrigues' formula; the angle is taken as radians, as the library
// takes it).
void mat_translate(float *m, const float *v) {
    for (int r = 0; r < 4; ++r) m[12 + r] = ((m[r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r];
}
void mat_scale(float *m, const float *v) {
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 4; ++r) m[4 * c + r] *= v[c];
}
void mat_rotate(float *m, float angle, const float *axis) {
    const float c = cos(angle), s = sin(angle);
    const float t[3] = {(1.0f - c) * axis[0], (1.0f - c) * axis[1], (1.0f - c) * axis[2]};
    float R[3][3];   // R[column][row]
    R[0][0] = c + t[0] * axis[0];
    R[0][1] = t[0] * axis[1] + s * axis[2];
    R[0][2] = t[0] * axis[2] - s * axis[1];
    R[1][0] = t[1] * axis[0] - s * axis[2];
    R[1][1] = c + t[1] * axis[1];
    R[1][2] = t[1] * axis[2] + s * axis[0];
    R[2][0] = t[2] * axis[0] + s * axis[1];
    R[2][1] = t[2] * axis[1] - s * axis[0];
    R[2][2] = c + t[2] * axis[2];
    float out[12];
    for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 4; ++r) out[4 * k + r] = (m[r] * R[k][0] + m[4 + r] * R[k][1]) + m[8 + r] * R[k][2];
    for (int i = 0; i < 12; ++i) m[i] = out[i];
}

// scene::find_model (libgi/scene.cpp:66-73): a relative path is looked up
// under each model path in order, else taken as given.
string find_model(const string &path) {
    if (!path.empty() && path[0] != '/')
        for (const string &p : modelpaths) {
            string cand = p.empty() ? path : (p.back() == '/' ? p + path : p + "/" + path);
            if (ifstream(cand).good()) return cand;
        }
    return path;
}

string material_name(hiprt_host_scene s, uint32_t m) {
    char buf[256];
    return hiprt_scene_material_name(s, m, buf, sizeof buf) == HIPRT_OK ? string(buf) : string();
}

string rest_of_line(istringstream &in) {
    string r;
    while (isspace(in.peek())) in.get();
    getline(in, r);
    return r;
}

#define error(x)                                                              \
    {                                                                         \
        cerr << "command " << uc.cmdid << " (" << command << "): " << x << endl; \
        return;                                                               \
    }
#define check_in_complete(x)                                                  \
    {                                                                         \
        if (in.bad() || in.fail() || !in.eof()) error(x);                     \
    }

bool write_image(const string &path) {
    const auto &fb = wf::rc->framebuffer;
    int w = wf::rc->w, h = wf::rc->h;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    bool pfm = path.size() > 4 && path.substr(path.size() - 4) == ".pfm";
    if (pfm) {
        fprintf(f, "PF\n%d %d\n-1.0\n", w, h);   // little endian, bottom row first
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) fwrite(&fb[4 * ((size_t)y * w + x)], 4, 3, f);
    } else {
        fprintf(f, "P6\n%d %d\n255\n", w, h);
        for (int y = h - 1; y >= 0; --y)
            for (int x = 0; x < w; ++x)
                for (int k = 0; k < 3; ++k) {
                    float v = fb[4 * ((size_t)y * w + x) + k];
                    v = v < 0 ? 0 : (v > 1 ? 1 : v);
                    fputc((int)lround(255.0 * pow(v, 1.0 / 2.2)), f);
                }
    }
    return fclose(f) == 0;
}

bool read3(istringstream &in, float *v) {
    in >> v[0] >> v[1] >> v[2];
    return !in.fail();
}

void run() {
    algo->prepare_data();
    algo->prepare_frame();
    while (algo->compute_sample(wf::rc->sppx, wf::rc->platform->timer)) {
    }
    algo->finalize_frame(wf::rc->platform->timer);
    if (!write_image(outfile)) cerr << "cannot write " << outfile << endl;
}

void eval(const string &line) {
    istringstream in(line);
    string command;
    in >> command;
    ++uc.cmdid;
    wf::render_context *rc = wf::rc;
    if (skip_block) {   // interaction.cpp:238-244
        if (line == "}") skip_block = false;
        return;
    }
    try {
        if (command.empty() || command[0] == '#' || command.substr(0, 2) == "//") {
            return;
        } else if (command == "disable") {
            string what;
            in >> what;
            check_in_complete("Syntax error, expecting only one feature to disable");
            cout << "DISABLING " << what << endl;
            disabled_features[what] = true;
        } else if (command == "with") {
            // interaction.cpp:263-297; the platforms this build has: hip
            string what, rest;
            in >> what;
            bool missing;
            if (what == "gl" || what == "cuda" || what == "optix") missing = true;
            else if (what == "hip") missing = disabled_features[what];
            else error("Syntax error: expected one of gl, cuda, optix, hip");
            in >> rest;
            check_in_complete("Syntax error, expecting only {");
            if (rest != "{") error("Syntax error, expecting {");
            skip_block = missing;
            if (skip_block) cout << "Skipping a block due to missing " << what << endl;
        } else if (command == "}") {
            check_in_complete("Syntax error: } must be on an otherwise empty line");
        } else if (command == "material") {
            // interaction.cpp:624-706
            string sub;
            in >> sub;
            hiprt_host_scene s = rc->scene.geometry;
            uint32_t n_mats = 0;
            if (s) hiprt_scene_materials(s, nullptr, &n_mats);
            if (in.eof() && sub == "list") {
                for (uint32_t m = 0; m < n_mats; ++m) cout << material_name(s, m) << endl;
                return;
            }
            if (in.fail()) error("Syntax error, requires material name, command and subsequent arguments");
            command = sub;   // later errors name the subcommand (interaction.cpp:632)
            if (sub == "select") {
                string name = rest_of_line(in);
                int found = -1;
                for (uint32_t m = 0; m < n_mats && found < 0; ++m)
                    if (material_name(s, m) == name) found = (int)m;
                if (found < 0) error("No material called '" << name << "'");
                selected_mat = found;
                return;
            }
            if (sub == "blacklist") return;
            if (selected_mat < 0 || !s || (uint32_t)selected_mat >= n_mats) error("No material selected");
            if (sub == "albedo" || sub == "emissive") {
                float v[3];
                if (!read3(in, v)) error("Expects a color triplet");
                check_in_complete("Expects a color triplet");
                if (hiprt_scene_set_material(s, (uint32_t)selected_mat, sub == "albedo" ? v : nullptr,
                                             sub == "emissive" ? v : nullptr) != HIPRT_OK)
                    error(hiprt_last_error());
                uc.scene_touched_at = uc.cmdid;
            } else if (sub == "roughness" || sub == "ior") {
                float x;
                in >> x;
                check_in_complete("Expects a floating point value");
                (sub == "roughness" ? mat_params[selected_mat].roughness : mat_params[selected_mat].ior) = x;
            } else if (sub == "brdf") {
                string name;
                in >> name;
                check_in_complete("Expected a single (no whitespace) string value");
                if (!known_brdf(name)) error("No such brdf defined: " << name);
                mat_params[selected_mat].brdf = name;
            } else if (sub == "texture") {
                error("Texture files are not decoded here; attach RGBA texels with hiprt_scene_set_texture");
            } else if (sub == "show") {
                check_in_complete("Does not take further arguments");
                vector<float> t(8ull * n_mats);
                hiprt_scene_materials(s, t.data(), &n_mats);
                const float *m = &t[8ull * selected_mat];
                const brdf_params &bp = mat_params[selected_mat];
                cout << "albedo     " << m[0] << " " << m[1] << " " << m[2] << endl;
                cout << "albedo-tex none" << endl;
                cout << "emissive   " << m[4] << " " << m[5] << " " << m[6] << endl;
                cout << "roughness  " << bp.roughness << endl;
                cout << "ior        " << bp.ior << endl;
            } else {
                error("Unknown subcommand");
            }
        } else if (command == "default-brdf") {
            // interaction.cpp:709-722
            string name;
            in >> name;
            check_in_complete("Expected a single (no whitespace) string value");
            if (brdfs_applied) cout << "Replacing default brdfs that has already been applied to some materials." << endl;
            if (!known_brdf(name)) error("No such brdf defined: " << name);
            default_brdf = name;
        } else if (command == "at" || command == "look" || command == "camup" || command == "up") {
            // interaction.cpp:301-350: without arguments they print the
            // vector; "up" sets the scene's up vector once, and the camera's
            // too unless "camup" set that
            float *cur = command == "at" ? rc->scene.cam_pos
                         : command == "look" ? rc->scene.cam_dir
                         : command == "camup" ? rc->scene.cam_up : scene_up;
            in.get();
            if (in.eof() && (command != "up" || scene_up_set)) {
                cout << cur[0] << " " << cur[1] << " " << cur[2] << endl;
                return;
            }
            if (command == "up" && scene_up_set) error("Cannot set scene up vector twice, did you mean camup?");
            float v[3];
            in >> v[0] >> v[1] >> v[2];
            check_in_complete("Syntax error, requires 3 numerical components");
            for (int k = 0; k < 3; ++k) cur[k] = v[k];
            if (command == "camup") cam_has_up = true;
            if (command == "up") {
                scene_up_set = true;
                if (!cam_has_up) {
                    for (int k = 0; k < 3; ++k) rc->scene.cam_up[k] = v[k];
                    cam_has_up = true;
                }
            }
        } else if (command == "skylight") {
            // interaction.cpp:754-777
            string sub;
            in >> sub;
            if (sub == "load") {
                float scale;
                string file;
                in >> file >> scale;
                check_in_complete("Syntax error: skylight load filename intensity-scale (note: filename is not allowed "
                                  "to contain spaces)");
                string path = find_model(file);
                if (!ifstream(path, ios::binary).good())
                    error("Cannot open file '" << path << "' for hdr floats texture.");
                have_sky = true;
                sky_scale = scale;
            } else if (sub == "scale") {
                float scale;
                in >> scale;
                check_in_complete("Syntax error: skylight scale new-scale-value");
                if (!have_sky) error("There is no skylight to scale");
                if (scale <= 0) error("Intensity scale has to be a value > 0");
                sky_scale = scale;
            } else {
                error("No such skylight subcommand");
            }
        } else if (command == "preview-offset") {
            // interaction.cpp:598-605 (no preview window here)
            int offset;
            in >> offset;
            check_in_complete("Syntax error, requires exactly one positive integral value");
            if (offset <= 0) error("Thge preview offset needs to be larger than zero");
        } else if (command == "omp") {
            // interaction.cpp:819-828: host threads; the tracing runs on the GPU
            string sub;
            in >> sub;
            if (sub != "on" && sub != "off") error("Syntax error: expected 'on' or 'off'");
        } else if (command == "denoise") {
            // interaction.cpp:573-590 on a build without OpenImageDenoise
            string option;
            in >> option;
            if (option == "on") cout << "Denoising not available: OpenImageDenoise library not found" << endl;
            else if (option != "off") error("Invalid option for denoise, valid are on/off");
        } else if (command == "fovy") {
            float f;
            in >> f;
            check_in_complete("Syntax error, requires one angle in degrees");
            rc->scene.fovy = f;
        } else if (command == "resolution") {
            int w, h;
            in >> w >> h;
            check_in_complete("Syntax error, requires 2 integral values");
            if (w <= 0 || h <= 0) error("Resolution must be positive");
            rc->w = w;
            rc->h = h;
            algo.reset();   // ray data is sized by the resolution
            algo_name.clear();
        } else if (command == "sppx") {
            int s;
            in >> s;
            check_in_complete("Syntax error, requires exactly one positive integral value");
            rc->sppx = s;
        } else if (command == "load") {
            // interaction.cpp:395-402: the model joins the scene under the
            // current model matrix, its materials named "<name>/..." when a
            // name is given
            string file, name;
            in >> file;
            if (!in.eof()) in >> name;
            check_in_complete("Syntax error, requires a file name (no spaces, sorry) and (optionally) a name");
            string path = find_model(file);
            if (hiprt_scene_add_obj(&rc->scene.geometry, path.c_str(), name.empty() ? nullptr : name.c_str(),
                                    modelmatrix) != HIPRT_OK)
                error(hiprt_last_error());
            brdfs_applied = true;
            uc.scene_touched_at = uc.cmdid;
        } else if (command == "modelmatrix") {
            // interaction.cpp:357-363: 16 values, column by column
            for (float &v : modelmatrix) in >> v;
            check_in_complete("Syntax error in model matrix");
        } else if (command == "modeltrafo") {
            // interaction.cpp:364-393
            string trafo;
            in >> trafo;
            if (trafo == "shift" || trafo == "translate") {
                float v[3];
                in >> v[0] >> v[1] >> v[2];
                check_in_complete("Too many components in shift transformation");
                mat_translate(modelmatrix, v);
            } else if (trafo == "rotate") {
                const float axes[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                string a;
                in >> a;
                int aid = 0;
                if (a == "x" || a == "X") aid = 0;
                else if (a == "y" || a == "Y") aid = 1;
                else if (a == "z" || a == "Z") aid = 2;
                else error("Invalid axis of rotation; " << a);
                float deg;
                in >> deg;
                check_in_complete("Rotation requires one angle (in degrees)");
                mat_rotate(modelmatrix, deg, axes[aid]);
            } else if (trafo == "scale") {
                float v[3];
                in >> v[0] >> v[1] >> v[2];
                check_in_complete("Too many components in scale transformation");
                mat_scale(modelmatrix, v);
            } else {
                error("unsupported trafo: " << trafo);
            }
        } else if (command == "modelpath") {
            // interaction.cpp:404-409, scene::add_modelpath (libgi/scene.cpp:44-53)
            string name;
            in >> name;
            check_in_complete("Syntax error, requires a path name (no spaces, sorry)");
            if (name == "~" || name.rfind("~/", 0) == 0) {
                const char *home = getenv("HOME");
                name = string(home ? home : "") + name.substr(1);
            }
            modelpaths.push_back(name);
        } else if (command == "generate") {
            int cfg;
            in >> cfg;
            check_in_complete("Syntax error, requires a workload number (1-3)");
            hiprt_host_scene s = nullptr;
            float cam[15];
            if (hiprt_scene_generate(cfg, &s, cam) != HIPRT_OK) error(hiprt_last_error());
            if (rc->scene.geometry) hiprt_scene_free(rc->scene.geometry);
            rc->scene.geometry = s;
            selected_mat = -1;
            mat_params.clear();
            for (int k = 0; k < 3; ++k) {
                rc->scene.cam_pos[k] = cam[k];
                rc->scene.cam_dir[k] = cam[3 + k];
                rc->scene.cam_up[k] = cam[6 + k];
            }
            rc->scene.fovy = cam[9];
            rc->scene.lights = {wf::pointlight{{cam[12], cam[13], cam[14]}, {400.0f, 400.0f, 400.0f}}};
            uc.scene_touched_at = uc.cmdid;
        } else if (command == "pointlight") {
            // pointlight [replace] pos x y z col r g b (interaction.cpp:723-753)
            string cmd;
            in >> cmd;
            bool replace = false;
            if (cmd == "replace") {
                replace = true;
                in >> cmd;
            }
            wf::pointlight pl;
            if (cmd != "pos" || !read3(in, pl.pos)) error("Syntax error: poinlight [replace] pos x y z col x y z");
            in >> cmd;
            if (cmd != "col" || !read3(in, pl.col)) error("Syntax error: poinlight [replace] pos x y z col x y z");
            check_in_complete("Syntax error: poinlight [replace] pos x y z col x y z");
            if (replace && !rc->scene.lights.empty()) rc->scene.lights[0] = pl;
            else rc->scene.lights.push_back(pl);
        } else if (command == "platform") {
            string name;
            in >> name;
            if (name.empty()) {
                cout << "platform: " << (rc->platform ? rc->platform->name : string("none")) << endl;
                return;
            }
            vector<string> args;
            string a;
            while (in >> a) args.push_back(a);
            algo.reset();
            algo_name.clear();
            delete rc->platform;
            rc->platform = nullptr;
            uc.valid_platform = false;
            if (name == "hip") {
                try {
                    rc->platform = new wf::hip::platform(args);
                } catch (const std::exception &e) {
                    error("The hip platform is not available: " << e.what());
                }
                uc.valid_platform = true;
            } else if (name == "none") {
                uc.valid_platform = true;
            } else {
                error("There is no platform called '" << name << "'");
            }
            uc.tracer_touched_at = uc.cmdid;
        } else if (command == "raytracer" || command == "bvh" || command == "triangles" ||
                   command == "tracer-option" || command == "incoherence") {
            if (!rc->platform) error("No platform selected, this command needs the hip platform");
            if (!rc->platform->interprete(command, in)) error("unknown command");
            if (command == "raytracer") uc.tracer_touched_at = uc.cmdid;
        } else if (command == "add-scene-step") {
            string name;
            while (isspace(in.peek())) in.get();
            getline(in, name);
            if (!rc->platform) error("No platform selected, steps are available only in wavefront mode");
            if (auto *s = rc->platform->step(name)) rc->platform->append_setup_step(s);
            else error("Platform " << rc->platform->name << " does not support step " << name);
        } else if (command == "commit") {
            if (!uc.valid_platform) error("Invalid platform");
            if (!rc->scene.geometry) error("There is no scene data to work with");
            if (!rc->platform) error("There is no ray traversal scheme to commit the scene data to");
            rc->platform->commit_scene(&rc->scene);
            uc.accel_touched_at = uc.cmdid;
        } else if (command == "algo") {
            string name;
            in >> name;
            check_in_complete("Syntax error, requires an algorithm name");
            auto a = wf::hip::make_algorithm(name);
            if (!a) error("There is no wavefront algorithm called '" << name << "'");
            algo = std::move(a);
            algo_name = name;
        } else if (command == "run") {
            if (!uc.valid_platform) error("Invalid platform");
            if (uc.scene_touched_at == 0 || uc.tracer_touched_at == 0 || uc.accel_touched_at == 0 || !algo)
                error("We have to have a scene loaded, a ray tracer set, an acceleration structure built and an "
                      "algorithm set prior to running");
            if (uc.accel_touched_at < uc.scene_touched_at) error("The current acceleration structure is out-dated");
            auto *pf = dynamic_cast<wf::hip::platform *>(rc->platform);
            if (!pf || !pf->ds) error("The committed scene has no acceleration structure (add-scene-step build accel struct)");
            run();
        } else if (command == "outfile") {
            string name;
            in >> name;
            check_in_complete("Syntax error, only accepts a single file name (no spaces, sorry)");
            outfile = name;
        } else if (command == "stats") {
            string sub;
            in >> sub;
            if (sub == "clear") wf::hip::clear_stats();
            else if (sub == "print") wf::hip::print_stats(cout);
            else error("No such stats subcommand");
        } else if (command == "echo") {
            string text;
            getline(in, text);
            cout << (text.empty() ? string(" ") : text.substr(text[0] == ' ' ? 1 : 0)) << endl;
        } else if (command == "quit" || command == "exit") {
            exit(0);
        } else if (algo && algo->interprete(command, in)) {
            // an algorithm's own command, e.g. direct-wf's "is uniform|cosine|light"
        } else {
            error("Unknown command");
        }
    } catch (const std::exception &e) {
        error(e.what());
    }
}

} // namespace

int main(int argc, char **argv) {
    wf::render_context ctx;
    wf::rc = &ctx;
    auto process = [](istream &in) {
        string line;
        while (getline(in, line)) eval(line);
    };
    if (argc < 2) {
        process(cin);
    } else {
        for (int i = 1; i < argc; ++i) {
            ifstream f(argv[i]);
            if (!f) {
                cerr << "cannot open script " << argv[i] << endl;
                return 1;
            }
            // models are also looked up beside the script while it runs
            // (interaction.cpp:129-137)
            string script = argv[i];
            size_t slash = script.find_last_of('/');
            modelpaths.push_back(slash == string::npos ? string() : script.substr(0, slash));
            process(f);
            modelpaths.pop_back();
        }
    }
    algo.reset();
    delete ctx.platform;
    ctx.platform = nullptr;
    return 0;
}
