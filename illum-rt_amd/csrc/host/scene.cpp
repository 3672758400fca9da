#include "scene.h"

#include "../core/rng.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <tuple>
#include <sstream>

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

// Material of an MTL "newmtl" block as the reference's scene::add derives it
// (libgi/scene.cpp:317-332): albedo = Kd if its luma exceeds 1e-4, else Ks,
// raised to the power 2.2; emission = Ke.
struct mtl_entry {
    vec3f kd{0, 0, 0}, ks{0, 0, 0}, ke{0, 0, 0};
    material resolve() const {
        vec3f a = luma(kd) > 1e-4f ? kd : ks;
        return material{vec3f{std::pow(a.x, 2.2f), std::pow(a.y, 2.2f), std::pow(a.z, 2.2f)}, ke};
    }
};

bool read_mtl(const std::string &path, std::vector<std::string> &names, std::vector<mtl_entry> &mats,
              std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open material library " + path;
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "newmtl") {
            std::string name;
            ls >> name;
            names.push_back(name);
            mats.emplace_back();
        } else if (!mats.empty() && (tag == "Kd" || tag == "Ks" || tag == "Ke")) {
            vec3f c{0, 0, 0};
            ls >> c.x >> c.y >> c.z;
            (tag == "Kd" ? mats.back().kd : tag == "Ks" ? mats.back().ks : mats.back().ke) = c;
        }
    }
    return true;
}

} // namespace

bool read_obj(const std::string &path, scene &out, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    out = scene{};
    std::vector<vec3f> normals;
    std::vector<float> uvs;                 // "vt" entries, (u, v) pairs
    std::vector<long> corner_vt;            // per triangle corner: vt index, -1 none
    std::vector<long> corner_vn;            // per triangle corner: vn index, -1 none
    std::vector<std::string> mtl_names;
    std::vector<mtl_entry> mtl;
    bool legacy = false;        // "usemtl m<id>" without a material library (write_obj's files)
    uint32_t cur_mat = 0;       // index into out.mats once resolved; see below
    int32_t cur_named = -1;     // current MTL material, -1 = none
    std::vector<int32_t> tri_named;
    std::string dir;
    if (size_t slash = path.find_last_of('/'); slash != std::string::npos) dir = path.substr(0, slash + 1);
    std::string line;
    size_t lineno = 0;
    auto where = [&]() { return path + ":" + std::to_string(lineno) + ": "; };
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v" || tag == "vn") {
            float x, y, z;
            if (!(ls >> x >> y >> z)) {
                err = where() + "malformed " + tag;
                return false;
            }
            (tag == "v" ? out.pos : normals).push_back(vec3f{x, y, z});
        } else if (tag == "vt") {
            float u, v = 0.0f;
            if (!(ls >> u)) {
                err = where() + "malformed vt";
                return false;
            }
            ls >> v;   // a 1-component coordinate has v = 0
            uvs.push_back(u);
            uvs.push_back(v);
        } else if (tag == "f") {
            std::vector<long> ids, nids, tids;
            std::string tok;
            while (ls >> tok) {
                // "a", "a/b", "a//c", "a/b/c" (1-based, negative = relative)
                long v = std::strtol(tok.c_str(), nullptr, 10), n = 0, tx = 0;
                size_t s1 = tok.find('/');
                if (s1 != std::string::npos) {
                    size_t s2 = tok.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < tok.size()) n = std::strtol(tok.c_str() + s2 + 1, nullptr, 10);
                    if (s1 + 1 < tok.size() && tok[s1 + 1] != '/') tx = std::strtol(tok.c_str() + s1 + 1, nullptr, 10);
                }
                if (v < 0) v = (long)out.pos.size() + v + 1;
                if (n < 0) n = (long)normals.size() + n + 1;
                if (uvs.empty()) tx = 0;   // a file without "vt" lines: corner coordinate indices are ignored
                if (tx < 0) tx = (long)(uvs.size() / 2) + tx + 1;
                if (tx < 0 || (size_t)tx > uvs.size() / 2) {
                    err = where() + "texture coordinate index out of range";
                    return false;
                }
                tids.push_back(tx - 1);   // -1: no texture coordinate
                if (v <= 0 || (size_t)v > out.pos.size()) {
                    err = where() + "face index out of range";
                    return false;
                }
                if (n < 0 || (size_t)n > normals.size()) {
                    err = where() + "normal index out of range";
                    return false;
                }
                ids.push_back(v - 1);
                nids.push_back(n - 1);   // -1: no normal
            }
            // fan triangulation; winding follows the shading normals when
            // every corner has one (libgi/scene.cpp:264-277)
            for (size_t k = 2; k < ids.size(); ++k) {
                uint32_t a = (uint32_t)ids[0], b = (uint32_t)ids[k - 1], c = (uint32_t)ids[k];
                long ta = tids[0], tb = tids[k - 1], tcn = tids[k];
                long na = nids[0], nb = nids[k - 1], nc = nids[k];
                if (na >= 0 && nb >= 0 && nc >= 0) {
                    vec3f pa = out.pos[a], pb = out.pos[b], pc = out.pos[c];
                    vec3f g = cross(pb - pa, pc - pa);
                    vec3f m = (normals[na] + normals[nb] + normals[nc]) * 0.333f;
                    if (!(dot(g, m) > 0)) {
                        std::swap(b, c);
                        std::swap(tb, tcn);
                        std::swap(nb, nc);
                    }
                }
                out.idx.push_back(a);
                out.idx.push_back(b);
                out.idx.push_back(c);
                corner_vt.push_back(ta);
                corner_vt.push_back(tb);
                corner_vt.push_back(tcn);
                corner_vn.push_back(na);
                corner_vn.push_back(nb);
                corner_vn.push_back(nc);
                out.mat.push_back(cur_mat);
                tri_named.push_back(cur_named);
            }
        } else if (tag == "mtllib") {
            std::string name;
            while (ls >> name)
                if (!read_mtl(dir + name, mtl_names, mtl, err)) return false;
        } else if (tag == "usemtl") {
            std::string name;
            ls >> name;
            cur_named = -1;
            for (size_t k = 0; k < mtl_names.size(); ++k)
                if (mtl_names[k] == name) cur_named = (int32_t)k;
            if (cur_named < 0 && name.size() > 1 && name[0] == 'm' &&
                name.find_first_not_of("0123456789", 1) == std::string::npos) {
                legacy = true;
                cur_mat = (uint32_t)std::strtoul(name.c_str() + 1, nullptr, 10);
            } else {
                cur_mat = 0;
            }
        }
    }
    bool with_vn = false;
    for (long n : corner_vn) with_vn = with_vn || n >= 0;
    if (!uvs.empty() || with_vn) {
        // Texture coordinates and normals belong to face corners (assimp
        // gives every corner its own vertex, libgi/scene.cpp:251-262): each
        // distinct (position, vt, vn) corner becomes a vertex of its own
        // after the "v" vertices, which keep (0, 0) and no normal.  With
        // normals in the file, a corner without one gets its face's normal,
        // on a vertex of its own.
        const size_t nv = out.pos.size();
        if (!uvs.empty()) out.tc.assign(2 * nv, 0.0f);
        if (with_vn) out.norm.assign(3 * nv, 0.0f);
        std::map<std::tuple<uint32_t, long, long>, uint32_t> split;
        for (size_t k = 0; k < out.idx.size(); ++k) {
            long t = uvs.empty() ? -1 : corner_vt[k];
            long n = with_vn ? corner_vn[k] : -1;
            if (t < 0 && !with_vn) continue;
            if (with_vn && n < 0) n = -2 - (long)(k / 3);   // this face's own normal
            auto key = std::make_tuple(out.idx[k], t, n);
            auto it = split.find(key);
            if (it == split.end()) {
                uint32_t id = (uint32_t)out.pos.size();
                out.pos.push_back(out.pos[out.idx[k]]);
                if (!uvs.empty()) {
                    out.tc.push_back(t >= 0 ? uvs[2 * t] : 0.0f);
                    out.tc.push_back(t >= 0 ? uvs[2 * t + 1] : 0.0f);
                }
                if (with_vn) {
                    vec3f nn;
                    if (n >= 0) {
                        nn = normals[n];
                    } else {
                        const size_t f = 3 * (k / 3);
                        vec3f pa = out.pos[out.idx[f]], pb = out.pos[out.idx[f + 1]], pc = out.pos[out.idx[f + 2]];
                        nn = normalize(cross(pb - pa, pc - pa));
                    }
                    out.norm.push_back(nn.x);
                    out.norm.push_back(nn.y);
                    out.norm.push_back(nn.z);
                }
                it = split.emplace(key, id).first;
            }
            out.idx[k] = it->second;
        }
    }
    if (!mtl.empty()) {
        // material 0: faces without a known material (default albedo 0.8);
        // 1 + k: the k-th material of the libraries, in definition order
        out.mats.assign(1, material{v3(0.8f), v3(0.0f)});
        for (const mtl_entry &e : mtl) out.mats.push_back(e.resolve());
        out.mat_names.assign(1, "default");
        out.mat_names.insert(out.mat_names.end(), mtl_names.begin(), mtl_names.end());
        for (size_t t = 0; t < out.mat.size(); ++t) out.mat[t] = tri_named[t] >= 0 ? (uint32_t)tri_named[t] + 1 : 0u;
        return true;
    }
    uint32_t max_mat = 0;
    for (uint32_t m : out.mat) max_mat = m > max_mat ? m : max_mat;
    out.mats.assign(max_mat + 1, material{v3(0.8f), v3(0.0f)});
    out.mat_names.clear();
    for (uint32_t m = 0; m <= max_mat; ++m) out.mat_names.push_back(legacy ? "m" + std::to_string(m) : "default");
    return true;
}

} // namespace rt
