// Host-side triangle scene: positions, index triples, per-triangle material,
// a small diffuse material table, plus the seeded synthetic generators and a
// minimal OBJ reader/writer.
//
// The reference keeps vertices as {pos, norm, tc} and triangles as
// {a, b, c, material_id} (libgi/rt.h:35-47); positions and indices take part
// in intersection, shading normals and texture coordinates in shading and
// the opacity test.
#pragma once

#include "../core/vec.h"

#include <cstdint>
#include <string>
#include <vector>

namespace rt {

struct material {
    vec3f albedo;
    vec3f emissive;
    int32_t tex = -1;   // albedo texture (index into scene::texs), -1: none
};

// An RGBA texture as libgi/scene.h:20-56 keeps it: w x h texels, row y at
// texel[y * w], nearest-texel lookup with wrap.  Only the alpha channel
// (opacity) takes part in tracing.
struct texture {
    uint32_t w = 0, h = 0;
    std::vector<float> rgba;   // 4 floats per texel
};

struct scene {
    std::vector<vec3f> pos;        // vertex positions
    std::vector<uint32_t> idx;     // 3 per triangle
    std::vector<uint32_t> mat;     // material id per triangle
    std::vector<material> mats;
    // material names (material::name of the reference), one per material, or
    // empty for scenes whose materials have none
    std::vector<std::string> mat_names;
    std::vector<float> tc;         // per vertex (u, v), empty when the scene has no texture coordinates
    // per vertex shading normal (x, y, z), vertex::norm; empty when the scene
    // has none, and then the geometric normal shades (the face normals the
    // reference's loader generates for meshes without normals)
    std::vector<float> norm;
    std::vector<texture> texs;

    uint32_t n_tris() const { return (uint32_t)(idx.size() / 3); }
    vec3f a(uint32_t t) const { return pos[idx[3 * t + 0]]; }
    vec3f b(uint32_t t) const { return pos[idx[3 * t + 1]]; }
    vec3f c(uint32_t t) const { return pos[idx[3 * t + 2]]; }
};

// Camera as in libgi/camera.h:7-25: fovy is the vertical field of view in
// degrees, near_h = tanf(pi*fovy/360), near_w = aspect*near_h.
struct camera {
    vec3f pos, dir, up;
    float fovy;
    int w, h;
};

// Configuration of a synthetic workload (scene + camera), by number:
//   1: 8,192 random triangles, camera (0.5,0.5,-1) -> +z, 60 deg, 640x480
//   2: interior hall, ~74k triangles, 1920x1080 camera inside
//   3: hall at 2x tessellation + debris, exactly 262,144 triangles, 1920x1080
// Config 4's scenes come from gen_random_tris with per-scene sizes.
struct workload {
    scene sc;
    camera cam;
    vec3f light;   // point light for the shadow-ray workload
};

// n triangles with centroids uniform in [0,1]^3 and vertices displaced from
// the centroid by N(0, sigma^2) per coordinate; PCG32 stream `seed`.
scene gen_random_tris(uint32_t n, float sigma, uint64_t seed);

// The procedural hall.  `dense` selects the config-3 tessellation; `debris`
// debris triangles (log-normal edge length, median 0.05) are appended.
scene gen_hall(bool dense, uint32_t debris, uint64_t seed);

workload make_workload(int config);

// Number of debris triangles that brings the dense hall to exactly 2^18.
uint32_t hall_debris_for_config3();

// OBJ: write_obj writes "v x y z" with %.9g (round-trips every float) and
// "f a b c" (1-based), one object, no normals, materials as "usemtl m<id>"
// groups in triangle order.
//
// read_obj loads a file as the reference's scene::add does through assimp's
// OBJ importer (libgi/scene.cpp:292-356, 222-289):
//  * triangle order and materials follow the importer's objects and meshes
//    ("o" re-opens an object, "g" opens one, "usemtl" splits meshes), which
//    the reference walks object by object, mesh by mesh;
//  * materials: 0 "DefaultMaterial", then each "mtllib" library's in
//    definition order, then names "usemtl" uses without a definition in
//    first-use order; Kd defaults to 0.6 grey, and albedo = Kd^2.2 (Ks^2.2
//    when Kd's luma is ~0), emission = Ke (libgi/scene.cpp:317-332);
//  * v, vt, vn, f ("a", "a/b", "a//c", "a/b/c", negative = relative); quads
//    split from their concave corner, larger polygons fanned;
//  * `trafo` (a column-major 4x4 model matrix, null for identity) moves the
//    positions, and its inverse transpose the normals (normalized), before a
//    triangle whose geometric normal disagrees with the mean of its corners'
//    normals has b and c swapped (libgi/scene.cpp:264-277);
//  * texture coordinates and shading normals belong to face corners, as
//    assimp gives every corner its own vertex: each distinct (v, vt, vn)
//    corner becomes a vertex of its own after the file's "v" vertices.  In a
//    file with normals, a corner without one gets its triangle's flat normal
//    (the normal assimp generates for meshes without normals).
bool write_obj(const scene &s, const std::string &path);
bool read_obj(const std::string &path, scene &out, std::string &err, const float *trafo = nullptr);

// Appends src to dst as scene::add appends a model: vertices, triangles and
// materials after dst's (material ids and texture indices offset), material
// names prefixed "<prefix>/" when prefix is non-empty.  When only one side
// has shading normals, the other gets flat ones (every corner a vertex of
// its own); texture coordinates missing on one side are (0, 0).
void append_scene(scene &dst, const scene &src, const std::string &prefix);

// read_obj(path, trafo) appended to dst with name as the prefix.
bool add_obj(scene &dst, const std::string &path, const std::string &name, const float *trafo, std::string &err);

} // namespace rt
