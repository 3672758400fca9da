#include "scene.h"

#include "../core/rng.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <array>
#include <tuple>
#include <unordered_map>
#include <sstream>
#include <string_view>

namespace rt {

namespace {

// Standard normals by the Box-Muller transform over PCG32 floats, computed in
// double and cached in pairs.  std::normal_distribution is avoided because its
// algorithm is implementation-defined.
struct normal_gen {
    pcg32 &rng;
    bool have = false;
    double spare = 0;
    explicit normal_gen(pcg32 &r) : rng(r) {}
    double operator()() {
        if (have) {
            have = false;
            return spare;
        }
        double u1 = 1.0 - (double)rng.next_float();   // (0, 1]
        double u2 = (double)rng.next_float();
        double r = std::sqrt(-2.0 * std::log(u1));
        double th = 2.0 * M_PI * u2;
        spare = r * std::sin(th);
        have = true;
        return r * std::cos(th);
    }
};

struct mesh_builder {
    scene &s;
    uint32_t vert(double x, double y, double z) {
        s.pos.push_back(vec3f{(float)x, (float)y, (float)z});
        return (uint32_t)s.pos.size() - 1;
    }
    void tri(uint32_t a, uint32_t b, uint32_t c, uint32_t m) {
        s.idx.push_back(a);
        s.idx.push_back(b);
        s.idx.push_back(c);
        s.mat.push_back(m);
    }
    // quad p00 p10 p11 p01 -> (p00,p10,p11), (p00,p11,p01)
    void quad(uint32_t p00, uint32_t p10, uint32_t p11, uint32_t p01, uint32_t m) {
        tri(p00, p10, p11, m);
        tri(p00, p11, p01, m);
    }
    // A planar grid origin + i/nu*eu + j/nv*ev with shared vertices.
    void grid(const double o[3], const double eu[3], const double ev[3], int nu, int nv, uint32_t m) {
        uint32_t base = (uint32_t)s.pos.size();
        for (int j = 0; j <= nv; ++j)
            for (int i = 0; i <= nu; ++i) {
                double fu = (double)i / nu, fv = (double)j / nv;
                vert(o[0] + fu * eu[0] + fv * ev[0], o[1] + fu * eu[1] + fv * ev[1], o[2] + fu * eu[2] + fv * ev[2]);
            }
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nu; ++i) {
                uint32_t p00 = base + j * (nu + 1) + i;
                quad(p00, p00 + 1, p00 + (nu + 1) + 1, p00 + (nu + 1), m);
            }
    }
};

void random_frame(pcg32 &rng, double n[3], double t[3], double b[3]) {
    // uniform direction on the sphere: z = 1 - 2u, phi = 2 pi v
    double z = 1.0 - 2.0 * rng.next_float();
    double phi = 2.0 * M_PI * rng.next_float();
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    n[0] = r * std::cos(phi);
    n[1] = r * std::sin(phi);
    n[2] = z;
    double h[3] = {0, 0, 0};
    h[(std::fabs(n[0]) < 0.6) ? 0 : 1] = 1.0;
    // t = normalize(h x n), b = n x t
    t[0] = h[1] * n[2] - h[2] * n[1];
    t[1] = h[2] * n[0] - h[0] * n[2];
    t[2] = h[0] * n[1] - h[1] * n[0];
    double l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 3; ++k) t[k] /= l;
    b[0] = n[1] * t[2] - n[2] * t[1];
    b[1] = n[2] * t[0] - n[0] * t[2];
    b[2] = n[0] * t[1] - n[1] * t[0];
}

constexpr double kHallX = 40.0, kHallY = 15.0, kHallZ = 15.0;

} // namespace

scene gen_random_tris(uint32_t n, float sigma, uint64_t seed) {
    scene s;
    s.mats.push_back(material{v3(0.8f), v3(0.0f)});
    pcg32 rng(seed);
    normal_gen normal(rng);
    s.pos.reserve(3 * (size_t)n);
    s.idx.reserve(3 * (size_t)n);
    s.mat.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = rng.next_float();
        for (int v = 0; v < 3; ++v) {
            double p[3];
            for (int k = 0; k < 3; ++k) p[k] = c[k] + (double)sigma * normal();
            s.pos.push_back(vec3f{(float)p[0], (float)p[1], (float)p[2]});
            s.idx.push_back(3 * i + v);
        }
    }
    return s;
}

// Materials of the hall: 0 walls, 1 floor, 2 columns, 3 arches (emissive),
// 4 clutter, 5 debris.
scene gen_hall(bool dense, uint32_t debris, uint64_t seed) {
    scene s;
    s.mats = {
        material{v3(0.70f), v3(0.0f)},
        material{v3(0.50f), v3(0.0f)},
        material{v3(0.60f), v3(0.0f)},
        material{v3(0.80f), vec3f{4.0f, 3.8f, 3.2f}},
        material{v3(0.40f), v3(0.0f)},
        material{v3(0.30f), v3(0.0f)},
    };
    mesh_builder mb{s};
    const int f = dense ? 2 : 1;   // room-surface grid density
    {
        // floor y=0 and ceiling y=15, walls x=0/x=40 and z=0/z=15
        double o0[3] = {0, 0, 0}, ex[3] = {kHallX, 0, 0}, ez[3] = {0, 0, kHallZ}, ey[3] = {0, kHallY, 0};
        double oc[3] = {0, kHallY, 0};
        double ox1[3] = {kHallX, 0, 0}, oz1[3] = {0, 0, kHallZ};
        mb.grid(o0, ez, ex, 15 * f, 40 * f, 1);    // floor
        mb.grid(oc, ex, ez, 40 * f, 15 * f, 0);    // ceiling
        mb.grid(o0, ey, ez, 15 * f, 15 * f, 0);    // wall x=0
        mb.grid(ox1, ez, ey, 15 * f, 15 * f, 0);   // wall x=40
        mb.grid(o0, ex, ey, 40 * f, 15 * f, 0);    // wall z=0
        mb.grid(oz1, ey, ex, 15 * f, 40 * f, 0);   // wall z=15
    }
    // 24 columns: 2 rows of 12 along x, radius 0.45, floor to ceiling.
    const int segs = dense ? 128 : 64, rings = 18;
    for (int row = 0; row < 2; ++row)
        for (int i = 0; i < 12; ++i) {
            double cx = kHallX * (i + 0.5) / 12.0, cz = row == 0 ? 4.0 : 11.0, r = 0.45;
            uint32_t base = (uint32_t)s.pos.size();
            for (int k = 0; k <= rings; ++k)
                for (int j = 0; j < segs; ++j) {
                    double a = 2.0 * M_PI * j / segs;
                    mb.vert(cx + r * std::cos(a), kHallY * k / rings, cz + r * std::sin(a));
                }
            for (int k = 0; k < rings; ++k)
                for (int j = 0; j < segs; ++j) {
                    uint32_t p00 = base + k * segs + j, p10 = base + k * segs + (j + 1) % segs;
                    mb.quad(p00, p10, p10 + segs, p00 + segs, 2);
                }
        }
    // 12 arched strips spanning the two column rows, emissive.
    const int arc = dense ? 128 : 64;
    for (int i = 0; i < 12; ++i) {
        double cx = kHallX * (i + 0.5) / 12.0, w = 0.6;
        uint32_t base = (uint32_t)s.pos.size();
        for (int j = 0; j <= arc; ++j) {
            double th = M_PI * j / arc;
            double y = 11.5 + 3.5 * std::sin(th), z = 7.5 - 3.5 * std::cos(th);
            for (int k = 0; k <= 2; ++k) mb.vert(cx - w / 2 + w * k / 2.0, y, z);
        }
        for (int j = 0; j < arc; ++j)
            for (int k = 0; k < 2; ++k) {
                uint32_t p00 = base + j * 3 + k;
                mb.quad(p00, p00 + 1, p00 + 3 + 1, p00 + 3, 3);
            }
    }
    pcg32 rng(seed);
    normal_gen normal(rng);
    // 5,000 small randomly oriented quads near the floor.
    for (int q = 0; q < 5000; ++q) {
        double c[3] = {0.5 + 39.0 * rng.next_float(), 0.05 + 1.95 * rng.next_float(), 0.5 + 14.0 * rng.next_float()};
        double sz = 0.1 + 0.3 * rng.next_float();
        double n[3], t[3], b[3];
        random_frame(rng, n, t, b);
        uint32_t v[4];
        const double su[4] = {-0.5, 0.5, 0.5, -0.5}, sv[4] = {-0.5, -0.5, 0.5, 0.5};
        for (int k = 0; k < 4; ++k)
            v[k] = mb.vert(c[0] + sz * (su[k] * t[0] + sv[k] * b[0]), c[1] + sz * (su[k] * t[1] + sv[k] * b[1]),
                           c[2] + sz * (su[k] * t[2] + sv[k] * b[2]));
        mb.quad(v[0], v[1], v[2], v[3], 4);
    }
    // Debris: log-normal edge length (median 0.05, sigma_ln 0.5), near-
    // equilateral, random orientation, centres uniform in the hall.
    for (uint32_t d = 0; d < debris; ++d) {
        double c[3] = {kHallX * rng.next_float(), kHallY * rng.next_float(), kHallZ * rng.next_float()};
        double len = 0.05 * std::exp(0.5 * normal());
        double n[3], t[3], b[3];
        random_frame(rng, n, t, b);
        double phi0 = 2.0 * M_PI * rng.next_float();
        uint32_t v[3];
        for (int k = 0; k < 3; ++k) {
            double a = phi0 + 2.0 * M_PI * k / 3.0 + 0.3 * (2.0 * rng.next_float() - 1.0);
            double r = len / std::sqrt(3.0);
            v[k] = mb.vert(c[0] + r * (std::cos(a) * t[0] + std::sin(a) * b[0]),
                           c[1] + r * (std::cos(a) * t[1] + std::sin(a) * b[1]),
                           c[2] + r * (std::cos(a) * t[2] + std::sin(a) * b[2]));
        }
        mb.tri(v[0], v[1], v[2], 5);
    }
    return s;
}

uint32_t hall_debris_for_config3() {
    // room 22,800 + columns 110,592 + arches 6,144 + clutter 10,000 = 149,536
    const uint32_t hall = 2 * (15 * 2 * 40 * 2) * 2 + 2 * (15 * 2 * 15 * 2) * 2 + 2 * (40 * 2 * 15 * 2) * 2 +
                          24 * 128 * 18 * 2 + 12 * 128 * 2 * 2 + 5000 * 2;
    return (1u << 18) - hall;
}

workload make_workload(int config) {
    workload w;
    if (config == 1) {
        w.sc = gen_random_tris(8192, 0.02f, 1);
        w.cam = camera{v3(0.5f, 0.5f, -1.0f), v3(0.0f, 0.0f, 1.0f), v3(0.0f, 1.0f, 0.0f), 60.0f, 640, 480};
        w.light = v3(0.5f, 2.0f, 0.5f);
    } else if (config == 2 || config == 3) {
        w.sc = config == 2 ? gen_hall(false, 0, 2) : gen_hall(true, hall_debris_for_config3(), 3);
        w.cam = camera{v3(1.5f, 1.7f, 7.5f), v3(1.0f, 0.0f, 0.0f), v3(0.0f, 1.0f, 0.0f), 60.0f, 1920, 1080};
        w.light = v3(20.0f, 13.5f, 7.5f);
    } else {
        std::abort();
    }
    return w;
}

bool write_obj(const scene &s, const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (const vec3f &p : s.pos) std::fprintf(f, "v %.9g %.9g %.9g\n", p.x, p.y, p.z);
    uint32_t cur = 0xffffffffu;
    for (uint32_t t = 0; t < s.n_tris(); ++t) {
        uint32_t m = s.mat.empty() ? 0 : s.mat[t];
        if (m != cur && s.mats.size() > 1) {
            std::fprintf(f, "usemtl m%u\n", m);
            cur = m;
        }
        std::fprintf(f, "f %u %u %u\n", s.idx[3 * t] + 1, s.idx[3 * t + 1] + 1, s.idx[3 * t + 2] + 1);
    }
    return std::fclose(f) == 0;
}

namespace {

float luma(vec3f c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// An MTL material as assimp's OBJ importer holds it (diffuse defaults to 0.6
// grey, specular and emission to black), resolved as the reference's
// scene::add derives its material (libgi/scene.cpp:317-332): albedo = Kd if
// its luma exceeds 1e-4, else Ks, raised to the power 2.2; emission = Ke.
struct mtl_entry {
    vec3f kd{0.6f, 0.6f, 0.6f}, ks{0, 0, 0}, ke{0, 0, 0};
    material resolve() const {
        vec3f a = luma(kd) > 1e-4f ? kd : ks;
        return material{vec3f{std::pow(a.x, 2.2f), std::pow(a.y, 2.2f), std::pow(a.z, 2.2f)}, ke};
    }
};

// The importer's material list: "DefaultMaterial" first, then library
// materials in definition order, then names "usemtl" uses without a
// definition, in first-use order (each with the default values).
struct mtl_list {
    std::vector<std::string> names{"DefaultMaterial"};
    std::vector<mtl_entry> mats{mtl_entry{}};
    int32_t find(const std::string &n) const {
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == n) return (int32_t)k;
        return -1;
    }
    int32_t find_or_add(const std::string &n) {
        int32_t k = find(n);
        if (k >= 0) return k;
        names.push_back(n);
        mats.emplace_back();
        return (int32_t)names.size() - 1;
    }
};

// Whitespace-separated fields of one line, read in place.
struct field_cursor {
    const char *p;
    void skip() {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\f' || *p == '\v') ++p;
    }
    bool token(const char *&b, const char *&e) {
        skip();
        if (!*p) return false;
        b = p;
        while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\f' && *p != '\v') ++p;
        e = p;
        return true;
    }
    bool number(float &x) {
        skip();
        if (!*p) return false;
        char *end = nullptr;
        const float v = std::strtof(p, &end);
        if (end == p) return false;
        x = v;
        p = end;
        return true;
    }
};

std::string trimmed(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\f\v"), b = s.find_last_not_of(" \t\r\f\v");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// What follows the keyword on a line, trimmed.
std::string rest_of(std::istringstream &ls) {
    std::string r;
    std::getline(ls, r);
    return trimmed(r);
}

// "newmtl" (the rest of the line names the material; a bare "newmtl" is the
// default material) opens a material, or re-opens one of that name; "Kd",
// "Ks", "Ke" set its colours (one value sets red only, as assimp reads it).
bool read_mtl(std::istream &in, mtl_list &lib) {
    std::string line;
    int32_t cur = -1;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "newmtl") {
            std::string name = rest_of(ls);
            cur = lib.find_or_add(name.empty() ? "DefaultMaterial" : name);
        } else if (cur >= 0 && (tag == "Kd" || tag == "Ks" || tag == "Ke")) {
            vec3f c{0, 0, 0};
            ls >> c.x >> c.y >> c.z;
            mtl_entry &e = lib.mats[cur];
            (tag == "Kd" ? e.kd : tag == "Ks" ? e.ks : e.ke) = c;
        }
    }
    return true;
}

// The object and mesh bookkeeping of assimp's OBJ parser, which decides the
// triangle order and materials the reference's loader sees: it walks the
// node tree, one node per object in creation order, each node's meshes in
// creation order, each mesh's faces in file order (libgi/scene.cpp:222-289).
//   o name   re-opens the object of that name (the current mesh stays
//            current) or creates one;
//   g name   a group name other than the active one creates an object;
//   usemtl   switches material: a new mesh when the current one already has
//            faces and a different material; a mesh without a material yet
//            takes this one (faces already in it included);
//   f        creates an object ("defaultobject") and mesh when none is open.
// Every new object opens a mesh with the current material.
struct obj_corner {
    long v, t, n;   // 0-based indices; t, n = -1 for none
};
struct obj_face {
    uint32_t first, n;   // corners [first, first + n)
};
struct obj_mesh {
    int32_t mat = -1;   // -1: none yet (the importer then uses material 0)
    std::vector<obj_face> faces;
};
struct obj_object {
    std::string name;
    std::vector<uint32_t> meshes;
};
struct obj_model {
    std::vector<obj_object> objects;
    std::vector<obj_mesh> meshes;
    int32_t cur_object = -1, cur_mesh = -1, cur_mat = -1;
    std::string active_group;
    void create_mesh() {
        meshes.emplace_back();
        cur_mesh = (int32_t)meshes.size() - 1;
        if (cur_object >= 0) objects[(size_t)cur_object].meshes.push_back((uint32_t)cur_mesh);
    }
    void create_object(const std::string &name) {
        objects.push_back(obj_object{name, {}});
        cur_object = (int32_t)objects.size() - 1;
        create_mesh();
        if (cur_mat >= 0) meshes[(size_t)cur_mesh].mat = cur_mat;
    }
    void use_material(const std::string &name, mtl_list &lib) {
        if (name.empty() || (cur_mat >= 0 && lib.names[(size_t)cur_mat] == name)) return;
        cur_mat = lib.find_or_add(name);
        if (cur_mesh < 0) {
            create_mesh();
        } else {
            const obj_mesh &m = meshes[(size_t)cur_mesh];
            if (m.mat >= 0 && m.mat != cur_mat && !m.faces.empty()) create_mesh();
        }
        meshes[(size_t)cur_mesh].mat = cur_mat;
    }
    void open_object(const std::string &name) {
        for (size_t k = 0; k < objects.size(); ++k)
            if (objects[k].name == name) {
                cur_object = (int32_t)k;
                return;
            }
        create_object(name);
    }
    void open_group(const std::string &name) {
        if (name == active_group) return;
        create_object(name);
        active_group = name;
    }
    void add_face(obj_face f) {
        if (cur_object < 0) create_object("defaultobject");
        if (cur_mesh < 0) create_mesh();
        meshes[(size_t)cur_mesh].faces.push_back(f);
    }
};

// float arithmetic of a 4x4 column-major matrix (m[4 * column + row]) and
// 3-vectors, summed in the order the reference's matrix library sums.
vec3f xform_point(const float *m, vec3f p) {
    float r[3];
    for (int i = 0; i < 3; ++i) r[i] = (m[i] * p.x + m[4 + i] * p.y) + (m[8 + i] * p.z + m[12 + i] * 1.0f);
    return vec3f{r[0], r[1], r[2]};
}
// transpose(inverse(upper 3x3 of m)), as nm[3 * column + row]: the adjugate
// over the determinant.
void normal_matrix(const float *m, float *nm) {
    auto a = [m](int c, int r) { return m[4 * c + r]; };
    const float c00 = a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2);
    const float c01 = a(0, 1) * a(2, 2) - a(2, 1) * a(0, 2);
    const float c02 = a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2);
    const float inv_det = 1.0f / (a(0, 0) * c00 - a(1, 0) * c01 + a(2, 0) * c02);
    float inv[9];   // inverse, inv[3 * column + row]
    inv[0] = c00 * inv_det;
    inv[3] = -(a(1, 0) * a(2, 2) - a(2, 0) * a(1, 2)) * inv_det;
    inv[6] = (a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1)) * inv_det;
    inv[1] = -c01 * inv_det;
    inv[4] = (a(0, 0) * a(2, 2) - a(2, 0) * a(0, 2)) * inv_det;
    inv[7] = -(a(0, 0) * a(2, 1) - a(2, 0) * a(0, 1)) * inv_det;
    inv[2] = c02 * inv_det;
    inv[5] = -(a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2)) * inv_det;
    inv[8] = (a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1)) * inv_det;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) nm[3 * c + r] = inv[3 * r + c];
}
vec3f xform_normal(const float *nm, vec3f n) {
    vec3f t{nm[0] * n.x + nm[3] * n.y + nm[6] * n.z, nm[1] * n.x + nm[4] * n.y + nm[7] * n.z,
            nm[2] * n.x + nm[5] * n.y + nm[8] * n.z};
    const float d = (t.x * t.x + t.y * t.y) + t.z * t.z;
    const float s = 1.0f / std::sqrt(d);   // a zero vector gives NaN, as the reference's normalize does
    return vec3f{t.x * s, t.y * s, t.z * s};
}
vec3f cross_ref(vec3f x, vec3f y) {
    return vec3f{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
float dot_ref(vec3f a, vec3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// The flat normal assimp generates for a face without normals: the cross
// product of its first two edges, normalized unless it is zero.
vec3f face_normal(vec3f p0, vec3f p1, vec3f p2) {
    vec3f e1 = p1 - p0, e2 = p2 - p0;
    vec3f c{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
    const float len = std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
    if (len > 0.0f) c = vec3f{c.x / len, c.y / len, c.z / len};
    return c;
}

// assimp's split of a quad: fanned from its concave corner (the corner whose
// two angles to the opposite diagonal sum past pi), else from corner 0.
int quad_start(const vec3f *q) {
    auto unit = [](vec3f v) {
        const float len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        if (len == 0.0f) return v;
        const float inv = 1.0f / len;
        return vec3f{v.x * inv, v.y * inv, v.z * inv};
    };
    auto dot3 = [](vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    for (int i = 0; i < 4; ++i) {
        const vec3f v = q[i];
        const vec3f left = unit(q[(i + 3) % 4] - v), diag = unit(q[(i + 2) % 4] - v),
                    right = unit(q[(i + 1) % 4] - v);
        const float angle = std::acos(dot3(left, diag)) + std::acos(dot3(right, diag));
        if (angle > 3.14159265358979323846f) return i;
    }
    return 0;
}

} // namespace

bool read_obj(const std::string &path, scene &out, std::string &err, const float *trafo) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    out = scene{};
    std::vector<vec3f> pos, normals;
    std::vector<float> uvs;   // "vt" entries, (u, v) pairs
    std::vector<obj_corner> corners;
    mtl_list lib;
    obj_model model;
    std::string dir;
    if (size_t slash = path.find_last_of('/'); slash != std::string::npos) dir = path.substr(0, slash + 1);
    std::string line;
    size_t lineno = 0;
    auto where = [&]() { return path + ":" + std::to_string(lineno) + ": "; };
    while (std::getline(in, line)) {
        ++lineno;
        // geometry lines are parsed in place (they are nearly all of a file);
        // the rest through a string stream
        field_cursor fc{line.c_str()};
        const char *tb, *te;
        if (!fc.token(tb, te)) continue;
        const std::string_view tag(tb, (size_t)(te - tb));
        if (tag == "v" || tag == "vn") {
            float x, y, z;
            if (!fc.number(x) || !fc.number(y) || !fc.number(z)) {
                err = where() + "malformed " + std::string(tag);
                return false;
            }
            (tag == "v" ? pos : normals).push_back(vec3f{x, y, z});
            continue;
        }
        if (tag == "vt") {
            float u, v = 0.0f;
            if (!fc.number(u)) {
                err = where() + "malformed vt";
                return false;
            }
            fc.number(v);   // a 1-component coordinate has v = 0
            uvs.push_back(u);
            uvs.push_back(v);
            continue;
        }
        if (tag == "f") {
            const uint32_t first = (uint32_t)corners.size();
            const char *b, *e;
            while (fc.token(b, e)) {
                // "a", "a/b", "a//c", "a/b/c" (1-based, negative = relative)
                char *q = nullptr;
                long v = std::strtol(b, &q, 10), n = 0, tx = 0;
                if (q < e && *q == '/') {
                    if (q + 1 < e && q[1] == '/') {
                        if (q + 2 < e) n = std::strtol(q + 2, nullptr, 10);
                    } else if (q + 1 < e) {
                        char *q2 = nullptr;
                        tx = std::strtol(q + 1, &q2, 10);
                        if (q2 < e && *q2 == '/' && q2 + 1 < e) n = std::strtol(q2 + 1, nullptr, 10);
                    }
                }
                if (v < 0) v = (long)pos.size() + v + 1;
                if (n < 0) n = (long)normals.size() + n + 1;
                if (uvs.empty()) tx = 0;   // a file without "vt" lines: corner coordinate indices are ignored
                if (tx < 0) tx = (long)(uvs.size() / 2) + tx + 1;
                if (tx < 0 || (size_t)tx > uvs.size() / 2) {
                    err = where() + "texture coordinate index out of range";
                    return false;
                }
                if (v <= 0 || (size_t)v > pos.size()) {
                    err = where() + "face index out of range";
                    return false;
                }
                if (n < 0 || (size_t)n > normals.size()) {
                    err = where() + "normal index out of range";
                    return false;
                }
                corners.push_back(obj_corner{v - 1, tx - 1, n - 1});
            }
            if (corners.size() > first) model.add_face(obj_face{first, (uint32_t)corners.size() - first});
            continue;
        }
        std::istringstream ls(fc.p);
        if (tag == "mtllib") {
            // the rest of the line names one library, beside the OBJ file; a
            // missing one falls back to the OBJ's name with "mtl" for "obj",
            // and without that the file loads with no library
            std::string name = rest_of(ls);
            if (name.empty()) continue;
            std::ifstream lf(dir + name);
            if (!lf) lf = std::ifstream(path.substr(0, path.size() >= 3 ? path.size() - 3 : 0) + "mtl");
            if (lf) read_mtl(lf, lib);
        } else if (tag == "usemtl") {
            model.use_material(rest_of(ls), lib);
        } else if (tag == "o") {
            std::string name;
            if (ls >> name) model.open_object(name);
        } else if (tag == "g") {
            model.open_group(rest_of(ls));
        }
    }

    float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float *m = trafo ? trafo : ident;
    float nm[9];
    normal_matrix(m, nm);
    std::vector<vec3f> tpos(pos.size()), tnorm(normals.size());
    for (size_t i = 0; i < pos.size(); ++i) tpos[i] = xform_point(m, pos[i]);
    for (size_t i = 0; i < normals.size(); ++i) tnorm[i] = xform_normal(nm, normals[i]);
    bool with_vn = false;
    for (const obj_corner &c : corners) with_vn = with_vn || c.n >= 0;

    // Triangles in node order; per triangle corner: (v, vt, normal key), the
    // key a "vn" index or -2 - t for triangle t's own (face) normal.
    std::vector<obj_corner> tri_corners;
    std::vector<vec3f> own_normal;   // per triangle, its transformed face normal
    std::vector<std::array<uint32_t, 3>> tris;   // one face's triangles (corner positions in the face)
    for (const obj_object &o : model.objects)
        for (uint32_t mid : o.meshes) {
            const obj_mesh &mesh = model.meshes[mid];
            const uint32_t mat = mesh.mat < 0 ? 0u : (uint32_t)mesh.mat;
            for (const obj_face &f : mesh.faces) {
                const obj_corner *c = &corners[f.first];
                // quads as assimp splits them, larger polygons as a fan
                // (assimp clips ears there: same triangles for convex
                // polygons only up to their order)
                tris.clear();
                if (f.n == 4) {
                    const vec3f q[4] = {pos[(size_t)c[0].v], pos[(size_t)c[1].v], pos[(size_t)c[2].v],
                                        pos[(size_t)c[3].v]};
                    const uint32_t s = (uint32_t)quad_start(q);
                    tris.push_back({s, (s + 1) % 4, (s + 2) % 4});
                    tris.push_back({s, (s + 2) % 4, (s + 3) % 4});
                } else {
                    for (uint32_t k = 2; k < f.n; ++k) tris.push_back({0, k - 1, k});
                }
                auto emit = [&](uint32_t i0, uint32_t i1, uint32_t i2) {
                    obj_corner a = c[i0], b = c[i1], cc = c[i2];
                    const vec3f fn = face_normal(pos[(size_t)a.v], pos[(size_t)b.v], pos[(size_t)cc.v]);
                    const vec3f own = xform_normal(nm, fn);
                    auto nrm = [&](const obj_corner &k) { return k.n >= 0 ? tnorm[(size_t)k.n] : own; };
                    const vec3f mean = (nrm(a) + nrm(b) + nrm(cc)) * 0.333f;
                    const vec3f pa = tpos[(size_t)a.v];
                    const vec3f g = cross_ref(tpos[(size_t)b.v] - pa, tpos[(size_t)cc.v] - pa);
                    if (!(dot_ref(g, mean) > 0.0f)) std::swap(b, cc);   // libgi/scene.cpp:271-277
                    const long t = (long)own_normal.size();
                    for (obj_corner k : {a, b, cc}) {
                        if (k.n < 0) k.n = -2 - t;
                        tri_corners.push_back(k);
                    }
                    own_normal.push_back(own);
                    out.mat.push_back(mat);
                };
                for (const auto &t3 : tris) emit(t3[0], t3[1], t3[2]);
            }
        }

    out.pos = tpos;
    out.idx.resize(tri_corners.size());
    for (size_t k = 0; k < tri_corners.size(); ++k) out.idx[k] = (uint32_t)tri_corners[k].v;
    if (!uvs.empty() || with_vn) {
        // Texture coordinates and normals belong to face corners (assimp
        // gives every corner its own vertex, libgi/scene.cpp:251-262): each
        // distinct (position, vt, normal) corner becomes a vertex of its own
        // after the "v" vertices, which keep (0, 0) and no normal.  With
        // normals in the file, a corner without one gets its triangle's flat
        // normal, on a vertex of its own.
        const size_t nv = out.pos.size();
        if (!uvs.empty()) out.tc.assign(2 * nv, 0.0f);
        if (with_vn) out.norm.assign(3 * nv, 0.0f);
        struct key_hash {
            size_t operator()(const std::tuple<long, long, long> &k) const {
                uint64_t h = (uint64_t)std::get<0>(k) * 0x9E3779B97F4A7C15ull;
                h ^= (uint64_t)std::get<1>(k) * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
                h ^= (uint64_t)std::get<2>(k) * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
                return (size_t)h;
            }
        };
        std::unordered_map<std::tuple<long, long, long>, uint32_t, key_hash> split;
        split.reserve(tri_corners.size());
        for (size_t k = 0; k < tri_corners.size(); ++k) {
            const obj_corner &c = tri_corners[k];
            const long t = uvs.empty() ? -1 : c.t;
            const long n = with_vn ? c.n : -1;
            if (t < 0 && !with_vn) continue;
            auto key = std::make_tuple(c.v, t, n);
            auto it = split.find(key);
            if (it == split.end()) {
                const uint32_t id = (uint32_t)out.pos.size();
                out.pos.push_back(tpos[(size_t)c.v]);
                if (!uvs.empty()) {
                    out.tc.push_back(t >= 0 ? uvs[2 * (size_t)t] : 0.0f);
                    out.tc.push_back(t >= 0 ? uvs[2 * (size_t)t + 1] : 0.0f);
                }
                if (with_vn) {
                    const vec3f nn = n >= 0 ? tnorm[(size_t)n] : own_normal[(size_t)(-2 - n)];
                    out.norm.push_back(nn.x);
                    out.norm.push_back(nn.y);
                    out.norm.push_back(nn.z);
                }
                it = split.emplace(key, id).first;
            }
            out.idx[k] = it->second;
        }
    }
    out.mats.clear();
    for (const mtl_entry &e : lib.mats) out.mats.push_back(e.resolve());
    out.mat_names = lib.names;
    return true;
}

namespace {

// Gives every triangle corner a vertex of its own carrying the triangle's
// flat normal (and the corner's texture coordinate), for a scene without
// shading normals that is merged with one that has them.
void give_flat_normals(scene &s) {
    const uint32_t nt = s.n_tris();
    std::vector<vec3f> pos(3 * (size_t)nt);
    std::vector<float> tc(s.tc.empty() ? 0 : 6 * (size_t)nt), norm(9 * (size_t)nt);
    std::vector<uint32_t> idx(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const vec3f n = face_normal(s.a(t), s.b(t), s.c(t));
        for (int k = 0; k < 3; ++k) {
            const size_t j = 3 * (size_t)t + k, v = s.idx[j];
            pos[j] = s.pos[v];
            if (!tc.empty()) {
                tc[2 * j] = s.tc[2 * v];
                tc[2 * j + 1] = s.tc[2 * v + 1];
            }
            norm[3 * j] = n.x;
            norm[3 * j + 1] = n.y;
            norm[3 * j + 2] = n.z;
            idx[j] = (uint32_t)j;
        }
    }
    s.pos = std::move(pos);
    s.tc = std::move(tc);
    s.norm = std::move(norm);
    s.idx = std::move(idx);
}

} // namespace

void append_scene(scene &dst, const scene &src, const std::string &prefix) {
    scene add = src;
    if (dst.norm.empty() != add.norm.empty()) {
        // the side without shading normals gets the flat ones the
        // reference's loader generates (zeros for vertices of no triangle)
        scene &lacking = dst.norm.empty() ? dst : add;
        if (lacking.n_tris() > 0) give_flat_normals(lacking);
        else lacking.norm.assign(3 * lacking.pos.size(), 0.0f);
    }
    const uint32_t base_v = (uint32_t)dst.pos.size(), base_m = (uint32_t)dst.mats.size();
    const int32_t base_tex = (int32_t)dst.texs.size();
    if (!add.tc.empty() && dst.tc.empty()) dst.tc.assign(2 * dst.pos.size(), 0.0f);
    if (add.tc.empty() && !dst.tc.empty()) add.tc.assign(2 * add.pos.size(), 0.0f);
    dst.pos.insert(dst.pos.end(), add.pos.begin(), add.pos.end());
    dst.tc.insert(dst.tc.end(), add.tc.begin(), add.tc.end());
    dst.norm.insert(dst.norm.end(), add.norm.begin(), add.norm.end());
    for (uint32_t v : add.idx) dst.idx.push_back(v + base_v);
    for (uint32_t t = 0; t < add.n_tris(); ++t) dst.mat.push_back((add.mat.empty() ? 0u : add.mat[t]) + base_m);
    const bool named = !dst.mat_names.empty() || !add.mat_names.empty();
    if (named) dst.mat_names.resize(dst.mats.size());
    for (size_t k = 0; k < add.mats.size(); ++k) {
        material m = add.mats[k];
        if (m.tex >= 0) m.tex += base_tex;
        dst.mats.push_back(m);
        if (named) {
            std::string n = k < add.mat_names.size() ? add.mat_names[k] : std::string();
            dst.mat_names.push_back(prefix.empty() ? n : prefix + "/" + n);
        }
    }
    dst.texs.insert(dst.texs.end(), add.texs.begin(), add.texs.end());
}

bool add_obj(scene &dst, const std::string &path, const std::string &name, const float *trafo, std::string &err) {
    scene s;
    if (!read_obj(path, s, err, trafo)) return false;
    append_scene(dst, s, name);
    return true;
}

} // namespace rt
