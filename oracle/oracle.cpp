// CPU oracle for the batch ray-tracer path.  TEST INFRASTRUCTURE ONLY: it is
// loaded by the tests, by __graft_entry__.smoke() and by bench.py's CPU
// baseline as the checker; the product never links or calls it.
//
// It restates, in plain C++ compiled with -ffp-contract=off (as the reference
// is built: g++ -O2 on x86-64, no FMA):
//   ray construction         libgi/rt.h:24-31        (id = 1/d, t_min 1e-4)
//   Shirley triangle test    libgi/intersect.h:19-70
//   slab test intersect4     libgi/intersect.h:230-263
//   BVH closest/any hit      rt/cpu/bvh2.cpp:72-189  (binary_bvh_tracer, also alpha_aware)
//   opacity                  libgi/rt.cpp:11,32-35, libgi/scene.h:28-36
//   brute force              rt/cpu/seq.cpp:9-30     (seq_tri_is)
//   batch adapter            rt/cpu/wavefront.cpp:53-70 (OpenMP over rows)
//   camera rays              libgi/camera.h:18-24, libgi/camera.cpp:44-61
//   cosine bounce rays       rt/cpu/bounce.cpp:41-72, libgi/sampling.h:13-36,
//                            libgi/util.h:212-233, libgi/rt.cpp:7-15
// plus checks the reference does not have: an FP64 intersection of a given
// ray/triangle pair and counters of nodes visited and triangles tested.
//
// The tree comes in the reference's own bbvh_node layout (56-byte records,
// rt/cpu/bvh.h:42-52) with its leaf-slot index, so the oracle walks exactly
// the array the reference's tracer would.  One deliberate deviation: the
// reference's traversal stack is a fixed uint32_t[25] without a bounds check
// (bvh2.cpp:76), undefined behaviour on trees deeper than 24 inner levels;
// here it grows, which matches the reference wherever the reference is defined.
//
// Parity pins (tests/test_oracle.py): brute force == BVH traversal on small
// scenes; closed-form hits (centroid, edges, t bounds, degenerate and
// edge-on triangles); slab-test special cases; FP64 agreement; camera rays
// against the pinhole closed form; cosine rays against the hemisphere and
// cos-theta distribution.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

struct v3 {
    float x, y, z;
};
inline v3 add(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 sub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 mul(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
inline v3 smul(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline v3 cross(v3 a, v3 b) { return v3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline v3 normalize(v3 v) { return mul(v, 1.0f / sqrtf(dot(v, v))); }

// diff_geom::ns (libgi/rt.cpp:7-15): the corners' normals (vn, 3 floats per
// vertex) interpolated at (beta, gamma) and normalized.
inline v3 shading_normal(const float *vn, const uint32_t *idx, uint32_t tri, float beta, float gamma) {
    auto nv = [&](uint32_t v) { return v3{vn[3 * v], vn[3 * v + 1], vn[3 * v + 2]}; };
    return normalize(add(add(smul(1.0f - beta - gamma, nv(idx[3 * tri])), smul(beta, nv(idx[3 * tri + 1]))),
                         smul(gamma, nv(idx[3 * tri + 2]))));
}

struct ray {
    v3 o, d, id;
    float t_min, t_max;
};
inline ray make_ray(v3 o, v3 d, float tmin, float tmax) {
    ray r;
    r.o = o;
    r.d = d;
    r.id = v3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.t_min = tmin;
    r.t_max = tmax;
    return r;
}

struct hit {
    float t, beta, gamma;
    uint32_t ref;
};

struct box {
    v3 min, max;
};
// bbvh_node of rt/cpu/bvh.h:42-52
struct node {
    box l, r;
    int32_t link_l, link_r;
    bool inner() const { return link_r >= 0; }
};
static_assert(sizeof(node) == 56, "bbvh_node layout");

struct scene_view {
    const float *pos;
    const uint32_t *idx;
    uint32_t n_tris;
    v3 p(uint32_t v) const { return v3{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]}; }
};

// libgi/intersect.h:19-70
bool intersect_tri(const scene_view &s, uint32_t t, const ray &ray, hit &info) {
    v3 pa = s.p(s.idx[3 * t]), pb = s.p(s.idx[3 * t + 1]), pc = s.p(s.idx[3 * t + 2]);
    const float a_x = pa.x, a_y = pa.y, a_z = pa.z;
    const float a = a_x - pb.x, b = a_y - pb.y, c = a_z - pb.z;
    const float d = a_x - pc.x, e = a_y - pc.y, f = a_z - pc.z;
    const float g = ray.d.x, h = ray.d.y, i = ray.d.z;
    const float j = a_x - ray.o.x, k = a_y - ray.o.y, l = a_z - ray.o.z;
    float common1 = e * i - h * f;
    float common2 = g * f - d * i;
    float common3 = d * h - e * g;
    float M = a * common1 + b * common2 + c * common3;
    float beta = j * common1 + k * common2 + l * common3;
    common1 = a * k - j * b;
    common2 = j * c - a * l;
    common3 = b * l - k * c;
    float gamma = i * common1 + h * common2 + g * common3;
    float tt = -(f * common1 + e * common2 + d * common3);
    beta /= M;
    gamma /= M;
    tt /= M;
    if (tt > ray.t_min && tt < ray.t_max)
        if (beta > 0 && gamma > 0 && beta + gamma <= 1) {
            info.t = tt;
            info.beta = beta;
            info.gamma = gamma;
            return true;
        }
    return false;
}

// Opacity of the alpha-aware tracers (rt/cpu/bvh2.cpp:109-119, 171-176,
// ALPHA_THRESHOLD 0.5 of rt/cpu/bvh.h:56): diff_geom::opacity
// (libgi/rt.cpp:32-35) is 1 without an albedo texture, else the alpha of
// texture2d::sample (libgi/scene.h:28-36) at the interpolated texture
// coordinate of libgi/rt.cpp:11.  Textures are RGBA float texels.
struct alpha_view {
    const float *tc;          // per vertex (u, v)
    const uint32_t *mat;      // per triangle material id (null: all 0)
    const int32_t *mat_tex;   // per material texture index, -1 none
    uint32_t n_mats;
    const uint32_t *tex;      // per texture (first texel, w, h)
    const float *texels;      // 4 floats per texel
};

float opacity(const alpha_view &a, const scene_view &s, uint32_t tri, float beta, float gamma) {
    uint32_t m = a.mat ? a.mat[tri] : 0u;
    int32_t t = m < a.n_mats ? a.mat_tex[m] : -1;
    if (t < 0) return 1.0f;
    const uint32_t *ia = s.idx + 3ull * tri;
    float w0 = 1.0f - beta - gamma;
    float u = w0 * a.tc[2 * ia[0]] + beta * a.tc[2 * ia[1]] + gamma * a.tc[2 * ia[2]];
    float v = w0 * a.tc[2 * ia[0] + 1] + beta * a.tc[2 * ia[1] + 1] + gamma * a.tc[2 * ia[2] + 1];
    uint32_t first = a.tex[3 * t], w = a.tex[3 * t + 1], h = a.tex[3 * t + 2];
    u = u - floor(u);
    v = v - floor(v);
    int x = (int)(u * w + 0.5f);
    int y = (int)(v * h + 0.5f);
    if (x == (int)w) x = 0;
    if (y == (int)h) y = 0;
    // Non-finite coordinates: the reference reads outside the texture; the
    // oracle and the kernels both take texel 0.
    if (!(x >= 0 && x < (int)w)) x = 0;
    if (!(y >= 0 && y < (int)h)) y = 0;
    return a.texels[4ull * (first + (uint64_t)y * w + x) + 3];
}

constexpr float kAlphaThreshold = 0.5f;

// libgi/intersect.h:230-263
bool intersect4(const box &bx, const ray &ray, float &is) {
    float idx = ray.id.x, idy = ray.id.y, idz = ray.id.z;
    float t1x_tmp = (bx.min.x - ray.o.x) * idx;
    float t2x_tmp = (bx.max.x - ray.o.x) * idx;
    float t1x = (t1x_tmp < t2x_tmp) ? t1x_tmp : t2x_tmp;
    float t2x = (t2x_tmp < t1x_tmp) ? t1x_tmp : t2x_tmp;
    float t1y_tmp = (bx.min.y - ray.o.y) * idy;
    float t2y_tmp = (bx.max.y - ray.o.y) * idy;
    float t1y = (t1y_tmp < t2y_tmp) ? t1y_tmp : t2y_tmp;
    float t2y = (t2y_tmp < t1y_tmp) ? t1y_tmp : t2y_tmp;
    float t1z_tmp = (bx.min.z - ray.o.z) * idz;
    float t2z_tmp = (bx.max.z - ray.o.z) * idz;
    float t1z = (t1z_tmp < t2z_tmp) ? t1z_tmp : t2z_tmp;
    float t2z = (t2z_tmp < t1z_tmp) ? t1z_tmp : t2z_tmp;
    float t1 = (t1x < t1y) ? t1y : t1x;
    t1 = (t1z < t1) ? t1 : t1z;
    float t2 = (t2x < t2y) ? t2x : t2y;
    t2 = (t2z < t2) ? t2z : t2;
    if (t1 > t2) return false;
    if (t2 < ray.t_min) return false;
    if (t1 > ray.t_max) return false;
    is = t1;
    return true;
}

struct tree {
    const node *nodes;
    uint32_t n_nodes, root;
    const uint32_t *index;
};

// Traversal stack: a fixed array, as the reference's uint32_t stack[25]
// (rt/cpu/bvh2.cpp:76), continued on the heap past 64 entries, where the
// reference's would already have overflowed.  No allocation per ray.
struct trav_stack {
    uint32_t fixed[64];
    std::vector<uint32_t> more;
    uint32_t n = 0;
    bool empty() const { return n == 0; }
    void push_back(uint32_t v) {
        if (n < 64) fixed[n] = v;
        else more.push_back(v);
        ++n;
    }
    uint32_t pop() {
        --n;
        if (n < 64) return fixed[n];
        uint32_t v = more.back();
        more.pop_back();
        return v;
    }
};

// binary_bvh_tracer::closest_hit, rt/cpu/bvh2.cpp:72-137.  counts[0] += inner
// nodes visited, counts[1] += triangles tested (when counts is non-null).
hit closest_hit(const scene_view &s, const tree &t, const ray &ray, uint32_t *counts,
                const alpha_view *av = nullptr) {
    hit closest{FLT_MAX, 0, 0, 0}, intersection{FLT_MAX, 0, 0, 0};
    if (t.n_nodes == 0) return closest;
    trav_stack stack;
    stack.push_back(t.root);
    uint32_t n_inner = 0, n_tri = 0;
    while (!stack.empty()) {
        const node &nd = t.nodes[stack.pop()];
        if (nd.inner()) {
            ++n_inner;
            float dist_l, dist_r;
            bool hit_l = intersect4(nd.l, ray, dist_l) && dist_l < closest.t;
            bool hit_r = intersect4(nd.r, ray, dist_r) && dist_r < closest.t;
            if (hit_l && hit_r) {
                if (dist_l < dist_r) {
                    stack.push_back((uint32_t)nd.link_r);
                    stack.push_back((uint32_t)nd.link_l);
                } else {
                    stack.push_back((uint32_t)nd.link_l);
                    stack.push_back((uint32_t)nd.link_r);
                }
            } else if (hit_l) {
                stack.push_back((uint32_t)nd.link_l);
            } else if (hit_r) {
                stack.push_back((uint32_t)nd.link_r);
            }
        } else {
            int32_t off = -nd.link_l, cnt = -nd.link_r;
            for (int32_t i = 0; i < cnt; ++i) {
                uint32_t tri = t.index[off + i];
                ++n_tri;
                if (intersect_tri(s, tri, ray, intersection)) {
                    if (av && !(opacity(*av, s, tri, intersection.beta, intersection.gamma) > kAlphaThreshold))
                        continue;
                    if (intersection.t < closest.t) {
                        closest = intersection;
                        closest.ref = tri;
                    }
                }
            }
        }
    }
    if (counts) {
        counts[0] += n_inner;
        counts[1] += n_tri;
    }
    return closest;
}

// binary_bvh_tracer::any_hit, rt/cpu/bvh2.cpp:140-189
bool any_hit(const scene_view &s, const tree &t, const ray &ray, const alpha_view *av = nullptr) {
    hit intersection{FLT_MAX, 0, 0, 0};
    if (t.n_nodes == 0) return false;
    trav_stack stack;
    stack.push_back(t.root);
    while (!stack.empty()) {
        const node &nd = t.nodes[stack.pop()];
        if (nd.inner()) {
            float dist_l, dist_r;
            bool hit_l = intersect4(nd.l, ray, dist_l);
            bool hit_r = intersect4(nd.r, ray, dist_r);
            if (hit_l && hit_r) {
                if (dist_l < dist_r) {
                    stack.push_back((uint32_t)nd.link_r);
                    stack.push_back((uint32_t)nd.link_l);
                } else {
                    stack.push_back((uint32_t)nd.link_l);
                    stack.push_back((uint32_t)nd.link_r);
                }
            } else if (hit_l) {
                stack.push_back((uint32_t)nd.link_l);
            } else if (hit_r) {
                stack.push_back((uint32_t)nd.link_r);
            }
        } else {
            int32_t off = -nd.link_l, cnt = -nd.link_r;
            for (int32_t i = 0; i < cnt; ++i) {
                uint32_t tri = t.index[off + i];
                if (intersect_tri(s, tri, ray, intersection) &&
                    (!av || opacity(*av, s, tri, intersection.beta, intersection.gamma) > kAlphaThreshold))
                    return true;
            }
        }
    }
    return false;
}

// seq_tri_is, rt/cpu/seq.cpp:9-30
hit seq_closest(const scene_view &s, const ray &ray) {
    hit closest{FLT_MAX, 0, 0, 0}, intersection{FLT_MAX, 0, 0, 0};
    for (uint32_t i = 0; i < s.n_tris; ++i)
        if (intersect_tri(s, i, ray, intersection) && intersection.t < closest.t) {
            closest = intersection;
            closest.ref = i;
        }
    return closest;
}
bool seq_any(const scene_view &s, const ray &ray) {
    hit intersection{FLT_MAX, 0, 0, 0};
    for (uint32_t i = 0; i < s.n_tris; ++i)
        if (intersect_tri(s, i, ray, intersection)) return true;
    return false;
}

struct rays_in {
    const float *p[8];   // ox oy oz tmin dx dy dz tmax
    ray get(uint32_t i) const {
        return make_ray(v3{p[0][i], p[1][i], p[2][i]}, v3{p[4][i], p[5][i], p[6][i]}, p[3][i], p[7][i]);
    }
};

int set_threads(int threads) {
    int n = threads > 0 ? threads : omp_get_max_threads();
    return n;
}

} // namespace

extern "C" {

int orc_max_threads(void) { return omp_get_max_threads(); }

// Batch closest hit over `n` rays in rows of `row` rays, the OpenMP schedule
// of batch_rt_adapter::compute_closest_hit.  counts (2 per ray) may be null.
void orc_trace_closest(const float *pos, const uint32_t *idx, uint32_t n_tris, const void *nodes, uint32_t n_nodes,
                       uint32_t root, const uint32_t *index, const float *const *ray_planes, uint32_t n, uint32_t row,
                       float *t, float *u, float *v, uint32_t *id, uint32_t *counts, int threads) {
    scene_view s{pos, idx, n_tris};
    tree tr{static_cast<const node *>(nodes), n_nodes, root, index};
    rays_in R;
    for (int k = 0; k < 8; ++k) R.p[k] = ray_planes[k];
    if (row == 0) row = n ? n : 1;
    int64_t rows = ((int64_t)n + row - 1) / row;
#pragma omp parallel for schedule(static) num_threads(set_threads(threads))
    for (int64_t y = 0; y < rows; ++y)
        for (uint32_t x = 0; x < row; ++x) {
            uint64_t i = (uint64_t)y * row + x;
            if (i >= n) break;
            uint32_t c[2] = {0, 0};
            hit h = closest_hit(s, tr, R.get((uint32_t)i), counts ? c : nullptr);
            t[i] = h.t;
            u[i] = h.t != FLT_MAX ? h.beta : 0.0f;
            v[i] = h.t != FLT_MAX ? h.gamma : 0.0f;
            id[i] = h.ref;
            if (counts) {
                counts[2 * i] = c[0];
                counts[2 * i + 1] = c[1];
            }
        }
}

// Batch any hit: occluded[i] = 1 when any triangle is hit in (t_min, t_max).
void orc_trace_any(const float *pos, const uint32_t *idx, uint32_t n_tris, const void *nodes, uint32_t n_nodes,
                   uint32_t root, const uint32_t *index, const float *const *ray_planes, uint32_t n, uint8_t *occluded,
                   int threads) {
    scene_view s{pos, idx, n_tris};
    tree tr{static_cast<const node *>(nodes), n_nodes, root, index};
    rays_in R;
    for (int k = 0; k < 8; ++k) R.p[k] = ray_planes[k];
#pragma omp parallel for schedule(static) num_threads(set_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; ++i) occluded[i] = any_hit(s, tr, R.get((uint32_t)i)) ? 1 : 0;
}

// Alpha-aware batch traces (binary_bvh_tracer<..., alpha_aware = true>):
// tc per vertex, mat per triangle (may be null), mat_tex per material (-1:
// untextured), tex = (first texel, w, h) per texture, texels RGBA floats.
void orc_trace_closest_alpha(const float *pos, const uint32_t *idx, uint32_t n_tris, const void *nodes,
                             uint32_t n_nodes, uint32_t root, const uint32_t *index, const float *tc,
                             const uint32_t *mat, const int32_t *mat_tex, uint32_t n_mats, const uint32_t *tex,
                             const float *texels, const float *const *ray_planes, uint32_t n, float *t, float *u,
                             float *v, uint32_t *id, int threads) {
    scene_view s{pos, idx, n_tris};
    tree tr{static_cast<const node *>(nodes), n_nodes, root, index};
    alpha_view av{tc, mat, mat_tex, n_mats, tex, texels};
    rays_in R;
    for (int k = 0; k < 8; ++k) R.p[k] = ray_planes[k];
#pragma omp parallel for schedule(dynamic, 64) num_threads(set_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        hit h = closest_hit(s, tr, R.get((uint32_t)i), nullptr, &av);
        t[i] = h.t;
        u[i] = h.t != FLT_MAX ? h.beta : 0.0f;
        v[i] = h.t != FLT_MAX ? h.gamma : 0.0f;
        id[i] = h.ref;
    }
}

void orc_trace_any_alpha(const float *pos, const uint32_t *idx, uint32_t n_tris, const void *nodes, uint32_t n_nodes,
                         uint32_t root, const uint32_t *index, const float *tc, const uint32_t *mat,
                         const int32_t *mat_tex, uint32_t n_mats, const uint32_t *tex, const float *texels,
                         const float *const *ray_planes, uint32_t n, uint8_t *occluded, int threads) {
    scene_view s{pos, idx, n_tris};
    tree tr{static_cast<const node *>(nodes), n_nodes, root, index};
    alpha_view av{tc, mat, mat_tex, n_mats, tex, texels};
    rays_in R;
    for (int k = 0; k < 8; ++k) R.p[k] = ray_planes[k];
#pragma omp parallel for schedule(static) num_threads(set_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; ++i) occluded[i] = any_hit(s, tr, R.get((uint32_t)i), &av) ? 1 : 0;
}

// Opacity of triangle tri[i] at barycentrics (beta[i], gamma[i]).
void orc_opacity(const float *pos, const uint32_t *idx, uint32_t n_tris, const float *tc, const uint32_t *mat,
                 const int32_t *mat_tex, uint32_t n_mats, const uint32_t *tex, const float *texels,
                 const uint32_t *tri, const float *beta, const float *gamma, uint32_t n, float *out) {
    scene_view s{pos, idx, n_tris};
    alpha_view av{tc, mat, mat_tex, n_mats, tex, texels};
    for (uint32_t i = 0; i < n; ++i) out[i] = opacity(av, s, tri[i], beta[i], gamma[i]);
}

// Brute force over all triangles.
void orc_seq_closest(const float *pos, const uint32_t *idx, uint32_t n_tris, const float *const *ray_planes, uint32_t n,
                     float *t, float *u, float *v, uint32_t *id, int threads) {
    scene_view s{pos, idx, n_tris};
    rays_in R;
    for (int k = 0; k < 8; ++k) R.p[k] = ray_planes[k];
#pragma omp parallel for schedule(dynamic, 64) num_threads(set_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        hit h = seq_closest(s, R.get((uint32_t)i));
        t[i] = h.t;
        u[i] = h.t != FLT_MAX ? h.beta : 0.0f;
        v[i] = h.t != FLT_MAX ? h.gamma : 0.0f;
        id[i] = h.ref;
    }
}

void orc_seq_any(const float *pos, const uint32_t *idx, uint32_t n_tris, const float *const *ray_planes, uint32_t n,
                 uint8_t *occluded, int threads) {
    scene_view s{pos, idx, n_tris};
    rays_in R;
    for (int k = 0; k < 8; ++k) R.p[k] = ray_planes[k];
#pragma omp parallel for schedule(dynamic, 64) num_threads(set_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; ++i) occluded[i] = seq_any(s, R.get((uint32_t)i)) ? 1 : 0;
}

// FP64 intersection of ray i with triangle tri[i]: Cramer's rule on the fp32
// inputs widened to double.  Writes t, beta, gamma (NaN when the system is
// singular) and inside[i] = 1 when beta >= 0, gamma >= 0, beta + gamma <= 1
// (closed, so edge hits count as inside).
void orc_fp64_intersect(const float *pos, const uint32_t *idx, const uint32_t *tri, const float *const *ray_planes,
                        uint32_t n, double *t, double *beta, double *gamma, uint8_t *inside) {
    for (uint32_t r = 0; r < n; ++r) {
        const float *const *P = ray_planes;
        uint32_t tr = tri[r];
        double A[3], B[3], C[3];
        for (int k = 0; k < 3; ++k) {
            A[k] = pos[3 * idx[3 * tr] + k];
            B[k] = pos[3 * idx[3 * tr + 1] + k];
            C[k] = pos[3 * idx[3 * tr + 2] + k];
        }
        double o[3] = {P[0][r], P[1][r], P[2][r]}, d[3] = {P[4][r], P[5][r], P[6][r]};
        // Solve o + t d = A + beta (B-A) + gamma (C-A):
        // [B-A, C-A, -d] (beta, gamma, t)^T = o - A
        double e1[3], e2[3], q[3], m[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = B[k] - A[k];
            e2[k] = C[k] - A[k];
            q[k] = o[k] - A[k];
            m[k] = -d[k];
        }
        auto det3 = [](const double *c0, const double *c1, const double *c2) {
            return c0[0] * (c1[1] * c2[2] - c1[2] * c2[1]) - c1[0] * (c0[1] * c2[2] - c0[2] * c2[1]) +
                   c2[0] * (c0[1] * c1[2] - c0[2] * c1[1]);
        };
        double D = det3(e1, e2, m);
        if (D == 0.0) {
            t[r] = beta[r] = gamma[r] = NAN;
            inside[r] = 0;
            continue;
        }
        double b = det3(q, e2, m) / D, g = det3(e1, q, m) / D, tt = det3(e1, e2, q) / D;
        t[r] = tt;
        beta[r] = b;
        gamma[r] = g;
        inside[r] = (b >= 0 && g >= 0 && b + g <= 1) ? 1 : 0;
    }
}

// ---- closest-hit adjoint checks (no reference counterpart) ----------------
//
// solve64: o + t d = A + beta (B-A) + gamma (C-A) by Cramer's rule in double.
static bool solve64(const double *A, const double *B, const double *C, const double *o, const double *d, double &t,
                    double &beta, double &gamma) {
    double e1[3], e2[3], q[3], m[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = B[k] - A[k];
        e2[k] = C[k] - A[k];
        q[k] = o[k] - A[k];
        m[k] = -d[k];
    }
    auto det3 = [](const double *c0, const double *c1, const double *c2) {
        return c0[0] * (c1[1] * c2[2] - c1[2] * c2[1]) - c1[0] * (c0[1] * c2[2] - c0[2] * c2[1]) +
               c2[0] * (c0[1] * c1[2] - c0[2] * c1[1]);
    };
    double D = det3(e1, e2, m);
    if (D == 0.0) return false;
    beta = det3(q, e2, m) / D;
    gamma = det3(e1, q, m) / D;
    t = det3(e1, e2, q) / D;
    return true;
}

static void load15(const float *pos, const uint32_t *idx, uint32_t tri, const float *const *P, uint32_t r,
                   double *x) {   // A B C o d
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) x[3 * v + k] = pos[3 * idx[3 * tri + v] + k];
    x[9] = P[0][r];
    x[10] = P[1][r];
    x[11] = P[2][r];
    x[12] = P[4][r];
    x[13] = P[5][r];
    x[14] = P[6][r];
}

// Analytic FP64 gradient of L = gt*t + gb*beta + gg*gamma with respect to
// (A, B, C, o, d), 15 doubles per ray, via lambda = M^-T g (M = [B-A, C-A, -d]).
void orc_adjoint_fp64(const float *pos, const uint32_t *idx, const uint32_t *tri, const float *const *ray_planes,
                      uint32_t n, const double *gt, const double *gb, const double *gg, double *out) {
    for (uint32_t r = 0; r < n; ++r) {
        double x[15];
        load15(pos, idx, tri[r], ray_planes, r, x);
        const double *A = x, *B = x + 3, *C = x + 6, *o = x + 9, *d = x + 12;
        double t, b, g;
        double *y = out + 15 * (size_t)r;
        if (!solve64(A, B, C, o, d, t, b, g)) {
            for (int k = 0; k < 15; ++k) y[k] = NAN;
            continue;
        }
        // M^T lambda = (gb, gg, gt): rows of M^T are e1, e2, -d
        double M[3][3];
        for (int k = 0; k < 3; ++k) {
            M[0][k] = B[k] - A[k];
            M[1][k] = C[k] - A[k];
            M[2][k] = -d[k];
        }
        double rhs[3] = {gb[r], gg[r], gt[r]}, lam[3];
        double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                     M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        for (int c = 0; c < 3; ++c) {   // Cramer on the rows of M (system M lam = rhs)
            double Mc[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Mc[i][j] = j == c ? rhs[i] : M[i][j];
            lam[c] = (Mc[0][0] * (Mc[1][1] * Mc[2][2] - Mc[1][2] * Mc[2][1]) -
                      Mc[0][1] * (Mc[1][0] * Mc[2][2] - Mc[1][2] * Mc[2][0]) +
                      Mc[0][2] * (Mc[1][0] * Mc[2][1] - Mc[1][1] * Mc[2][0])) /
                     det;
        }
        for (int k = 0; k < 3; ++k) {
            y[k] = -(1.0 - b - g) * lam[k];
            y[3 + k] = -b * lam[k];
            y[6 + k] = -g * lam[k];
            y[9 + k] = lam[k];
            y[12 + k] = t * lam[k];
        }
    }
}

// Central differences of the same L with step h on each of the 15 inputs.
void orc_adjoint_fd(const float *pos, const uint32_t *idx, const uint32_t *tri, const float *const *ray_planes,
                    uint32_t n, const double *gt, const double *gb, const double *gg, double h, double *out) {
    for (uint32_t r = 0; r < n; ++r) {
        double x[15];
        load15(pos, idx, tri[r], ray_planes, r, x);
        double *y = out + 15 * (size_t)r;
        for (int k = 0; k < 15; ++k) {
            double v[2];
            for (int s = 0; s < 2; ++s) {
                double xp[15];
                for (int j = 0; j < 15; ++j) xp[j] = x[j];
                xp[k] += s == 0 ? h : -h;
                double t, b, g;
                if (!solve64(xp, xp + 3, xp + 6, xp + 9, xp + 12, t, b, g)) {
                    v[s] = NAN;
                    continue;
                }
                v[s] = gt[r] * t + gb[r] * b + gg[r] * g;
            }
            y[k] = (v[0] - v[1]) / (2 * h);
        }
    }
}

// Camera rays of libgi/camera.cpp:44-61 for a w x h image, t_min 1e-4,
// t_max FLT_MAX; out planes as ray_planes.  Sub-pixel offsets ox[i], oy[i]
// per pixel (both null: offset 0, the pixel centre).
void orc_camera_rays(const float *pos, const float *dir, const float *up, float fovy, int w, int h, const float *ox,
                     const float *oy, float *const *out) {
    v3 P{pos[0], pos[1], pos[2]}, D{dir[0], dir[1], dir[2]}, Up{up[0], up[1], up[2]};
    float aspect = float(w) / h;
    float near_h = tanf(float(M_PI) * fovy * 0.5f / 180.0f);
    float near_w = aspect * near_h;
    v3 U = normalize(cross(D, Up));
    v3 V = normalize(cross(U, D));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            float offx = ox ? ox[i] : 0.0f, offy = oy ? oy[i] : 0.0f;
            float uu = ((float)x + 0.5f + offx) / (float)w * 2.0f - 1.0f;
            float vv = ((float)y + 0.5f + offy) / (float)h * 2.0f - 1.0f;
            uu = near_w * uu;
            vv = near_h * vv;
            v3 d = normalize(add(add(D, mul(U, uu)), mul(V, vv)));
            out[0][i] = P.x;
            out[1][i] = P.y;
            out[2][i] = P.z;
            out[3][i] = 1e-4f;
            out[4][i] = d.x;
            out[5][i] = d.y;
            out[6][i] = d.z;
            out[7][i] = FLT_MAX;
        }
}

// Cosine-weighted bounce ray from hit (t, beta, gamma, id) of incoming ray i,
// driven by the given uniform numbers xi0/xi1 (rt/cpu/bounce.cpp:41-72): the
// normals flipped to the ray's side by the shading normal
// (flip_normals_to_ray, libgi/util.h:226-232; vn: 3 floats per vertex
// interpolated as diff_geom::ns, null: the geometric normal), the sample
// aligned to the geometric normal.  Rays without a hit get the reference's
// disabled ray: o = 0, d = (1,1,1), t_max = -FLT_MAX.
void orc_cosine_rays(const float *pos, const uint32_t *idx, const float *vn, const float *const *in_planes,
                     const float *hit_t, const float *hit_u, const float *hit_v, const uint32_t *hit_id,
                     const float *xi0, const float *xi1, uint32_t n, float *const *out) {
    const float pi = float(M_PI);
    for (uint32_t i = 0; i < n; ++i) {
        float o[3] = {0, 0, 0}, d[3] = {1, 1, 1}, tmax = -FLT_MAX;
        if (hit_t[i] != FLT_MAX) {
            uint32_t tr = hit_id[i];
            v3 A{pos[3 * idx[3 * tr]], pos[3 * idx[3 * tr] + 1], pos[3 * idx[3 * tr] + 2]};
            v3 B{pos[3 * idx[3 * tr + 1]], pos[3 * idx[3 * tr + 1] + 1], pos[3 * idx[3 * tr + 1] + 2]};
            v3 C{pos[3 * idx[3 * tr + 2]], pos[3 * idx[3 * tr + 2] + 1], pos[3 * idx[3 * tr + 2] + 2]};
            float beta = hit_u[i], gamma = hit_v[i];
            v3 x = add(add(smul(1.0f - beta - gamma, A), smul(beta, B)), smul(gamma, C));
            v3 ng = normalize(cross(sub(B, A), sub(C, A)));
            v3 ns = vn ? shading_normal(vn, idx, tr, beta, gamma) : ng;
            v3 rd{in_planes[4][i], in_planes[5][i], in_planes[6][i]};
            if (dot(rd, ns) > 0) ng = v3{-ng.x, -ng.y, -ng.z};
            // uniform_sample_disk + cosine_sample_hemisphere
            float r = sqrtf(xi0[i]);
            float theta = 2 * pi * xi1[i];
            float dx = r * cosf(theta), dy = r * sinf(theta);
            float z = 1 - dx * dx - dy * dy;
            v3 s{dx, dy, z > 0 ? sqrtf(z) : 0.f};
            // align(s, ng)
            float sg = copysignf(1.f, ng.z);
            v3 w{s.x, s.y, s.z * sg};
            v3 hv{ng.x, ng.y, ng.z + sg};
            float k = dot(w, hv) / (1.f + (ng.z < 0 ? -ng.z : ng.z));
            v3 wi = sub(smul(k, hv), w);
            o[0] = x.x;
            o[1] = x.y;
            o[2] = x.z;
            d[0] = wi.x;
            d[1] = wi.y;
            d[2] = wi.z;
            tmax = FLT_MAX;
        }
        out[0][i] = o[0];
        out[1][i] = o[1];
        out[2][i] = o[2];
        out[3][i] = 1e-4f;
        out[4][i] = d[0];
        out[5][i] = d[1];
        out[6][i] = d[2];
        out[7][i] = tmax;
    }
}

} // extern "C"

namespace {

// ---- wavefront path tracer of the workload --------------------------------
//
// PCG32 of the reference's rng_pcg (libgi/random.h:59-96, PBRT-v3's PCG32):
// state transition, XSH-RR output, stream selection, float mapping
// min(1 - 2^-24, u * 2^-32).
struct pcg32 {
    uint64_t state, inc;
    explicit pcg32(uint64_t initseq, uint64_t initstate = 0x853c49e6748fea9bULL) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_u32();
        state += initstate;
        next_u32();
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float next_float() {
        float f = (float)next_u32() * 0x1p-32f;
        return f < 0x1.fffffep-1f ? f : 0x1.fffffep-1f;
    }
};

// Stream of one (frame seed, path, bounce): SplitMix64's finaliser over
// seed * golden + (path << 8) + bounce + 1.  This is the workload's stream
// layout (the reference's wavefront steps draw from per-thread generators,
// which no GPU schedule can reproduce); the device kernels use the same.
uint64_t stream_of(uint64_t seed, uint64_t path, uint64_t bounce) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (path << 8) + bounce + 1;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// camera::update_frustum + cam_ray (libgi/camera.h:18-24, camera.cpp:44-61).
struct pinhole {
    v3 P, D, U, V;
    float near_w, near_h;
    int w, h;
    pinhole(const float *cam) {
        P = v3{cam[0], cam[1], cam[2]};
        D = v3{cam[3], cam[4], cam[5]};
        v3 Up{cam[6], cam[7], cam[8]};
        w = (int)cam[10];
        h = (int)cam[11];
        float aspect = float(w) / h;
        near_h = tanf(float(M_PI) * cam[9] * 0.5f / 180.0f);
        near_w = aspect * near_h;
        U = normalize(cross(D, Up));
        V = normalize(cross(U, D));
    }
    ray at(int x, int y, float ox, float oy) const {
        float uu = ((float)x + 0.5f + ox) / (float)w * 2.0f - 1.0f;
        float vv = ((float)y + 0.5f + oy) / (float)h * 2.0f - 1.0f;
        uu = near_w * uu;
        vv = near_h * vv;
        return make_ray(P, normalize(add(add(D, mul(U, uu)), mul(V, vv))), 1e-4f, FLT_MAX);
    }
};

// Cosine-weighted direction about the geometric normal flipped to the side
// the ray came from: sample_cos_weighted_dir (rt/cpu/bounce.cpp:41-72) with
// flip_normals_to_ray (libgi/util.h:226-233), uniform_sample_disk +
// cosine_sample_hemisphere (libgi/sampling.h:13-36) and align
// (libgi/util.h:212-218).  ns is the shading normal that decides the flip.
v3 cosine_dir(v3 ng, v3 ns, v3 in_d, float xi0, float xi1) {
    const float pi = float(M_PI);
    if (dot(in_d, ns) > 0) ng = v3{-ng.x, -ng.y, -ng.z};
    float r = sqrtf(xi0);
    float theta = 2 * pi * xi1;
    float dx = r * cosf(theta), dy = r * sinf(theta);
    float z = 1 - dx * dx - dy * dy;
    v3 s{dx, dy, z > 0 ? sqrtf(z) : 0.f};
    float sg = copysignf(1.f, ng.z);
    v3 w{s.x, s.y, s.z * sg};
    v3 hv{ng.x, ng.y, ng.z + sg};
    float k = dot(w, hv) / (1.f + (ng.z < 0 ? -ng.z : ng.z));
    return sub(smul(k, hv), w);
}

} // namespace

extern "C" {

// First n floats (next_float) of the PCG32 generator seeded with
// (initstate, initseq), for known-answer checks; raw != 0 writes next_u32
// values (as uint32 bits) instead.
void orc_pcg_sequence(uint64_t initstate, uint64_t initseq, uint32_t n, int raw, uint32_t *out) {
    pcg32 g(initseq, initstate);
    for (uint32_t i = 0; i < n; ++i) {
        if (raw) {
            out[i] = g.next_u32();
        } else {
            float f = g.next_float();
            std::memcpy(&out[i], &f, 4);
        }
    }
}

// The workload's diffuse path tracer: the control flow of simple_pt::path
// (gi/pt.cpp:72-115) with this workload's fixed choices.  Per frame seed s
// and pixel p (path p of that frame), from the camera ray of p -- through the
// pixel centre, or with jitter through the offset (xi0 - 0.5, xi1 - 0.5) of
// the first two uniforms of PCG32 stream (s, p, 255), as
// simple_pt::sample_pixel offsets every sample (gi/pt.cpp:36):
//   level L = 0 .. max_depth: closest hit of the current ray (the BVH
//   traversal above); a miss ends the path (no sky light, gi/pt.cpp:79-83);
//   an emissive material adds throughput * emission and ends the path
//   (gi/pt.cpp:86-89) -- or, with emitters_reflect, adds it and goes on;
//   at L = max_depth the path ends; otherwise two uniforms from the PCG32
//   stream (s, p, L) pick a cosine-weighted bounce (the wavefront step of
//   rt/cpu/bounce.cpp:41-72: origin = hit point, t_min 1e-4); when
//   L + 1 >= rr_start a third uniform xi keeps the path iff xi < rr_survive
//   (fixed-probability roulette instead of the luma-based one of
//   gi/pt.cpp:102-109) and scales by 1/rr_survive; the throughput is
//   multiplied by the albedo (f * cos / pdf of the cosine-sampled
//   Lambertian, albedo/pi * cos / (cos/pi)), and a zero throughput ends the
//   path (gi/pt.cpp:110-111).
// image (w*h*3) accumulates each pixel's path radiance frame by frame, in
// the order of frame_seeds; level_rays[L] (max_depth + 1 entries, may be
// null) accumulates the rays traced at level L; sig (w*h, may be null)
// receives, per pixel, a hash of the last frame's hit triangle ids and
// levels.  mats: 8 floats per material (albedo rgb, 0, emission rgb, 0);
// mat: material per triangle.  vn (3 floats per vertex, may be null):
// shading normals interpolated at the hit to decide the normal flip, as
// flip_normals_to_ray does with diff_geom::ns; null uses the geometric normal.
void orc_path_trace(const float *pos, const uint32_t *idx, uint32_t n_tris, const uint32_t *mat, const float *mats,
                    const float *vn, const void *nodes, uint32_t n_nodes, uint32_t root, const uint32_t *index,
                    const float *cam, int max_depth, int rr_start, float rr_survive, int emitters_reflect,
                    int jitter, const uint64_t *frame_seeds, uint32_t n_frames, float *image, uint64_t *level_rays,
                    uint64_t *sig, int threads) {
    scene_view s{pos, idx, n_tris};
    tree tr{static_cast<const node *>(nodes), n_nodes, root, index};
    pinhole ph(cam);
    const int64_t n_px = (int64_t)ph.w * ph.h;
    const int levels = max_depth + 1;
    std::vector<uint64_t> counts((size_t)levels, 0);
#pragma omp parallel num_threads(set_threads(threads))
    {
        std::vector<uint64_t> mine((size_t)levels, 0);
#pragma omp for schedule(dynamic, 256)
        for (int64_t p = 0; p < n_px; ++p) {
            int x = (int)(p % ph.w), y = (int)(p / ph.w);
            uint64_t h = 0;
            for (uint32_t f = 0; f < n_frames; ++f) {
                float ox = 0.0f, oy = 0.0f;
                if (jitter) {
                    pcg32 g(stream_of(frame_seeds[f], (uint64_t)p, 255));
                    ox = g.next_float() - 0.5f;
                    oy = g.next_float() - 0.5f;
                }
                ray r = ph.at(x, y, ox, oy);
                v3 T{1.0f, 1.0f, 1.0f};
                h = 1469598103934665603ULL;
                for (int L = 0; L <= max_depth; ++L) {
                    ++mine[(size_t)L];
                    hit c = closest_hit(s, tr, r, nullptr);
                    h = (h ^ (c.t == FLT_MAX ? 0xFFFFFFFFull : (uint64_t)c.ref)) * 1099511628211ULL;
                    h = (h ^ (uint64_t)L) * 1099511628211ULL;
                    if (c.t == FLT_MAX) break;
                    uint32_t tri = c.ref;
                    const float *m = mats + 8 * (size_t)(mat ? mat[tri] : 0u);
                    v3 alb{m[0], m[1], m[2]}, emi{m[4], m[5], m[6]};
                    if (emi.x != 0.0f || emi.y != 0.0f || emi.z != 0.0f) {
                        float *px = image + 3 * p;
                        px[0] += T.x * emi.x;
                        px[1] += T.y * emi.y;
                        px[2] += T.z * emi.z;
                        if (!emitters_reflect) break;
                    }
                    if (L >= max_depth) break;
                    pcg32 g(stream_of(frame_seeds[f], (uint64_t)p, (uint64_t)L));
                    float xi0 = g.next_float(), xi1 = g.next_float();
                    float scale = 1.0f;
                    if (L + 1 >= rr_start) {
                        float xi = g.next_float();
                        if (!(xi < rr_survive)) break;
                        scale = 1.0f / rr_survive;
                    }
                    v3 A = s.p(idx[3 * tri]), B = s.p(idx[3 * tri + 1]), C = s.p(idx[3 * tri + 2]);
                    float beta = c.beta, gamma = c.gamma;
                    v3 xh = add(add(smul(1.0f - beta - gamma, A), smul(beta, B)), smul(gamma, C));
                    v3 ng = normalize(cross(sub(B, A), sub(C, A)));
                    v3 ns = ng;
                    if (vn) {
                        auto nv = [&](uint32_t v) { return v3{vn[3 * v], vn[3 * v + 1], vn[3 * v + 2]}; };
                        ns = normalize(add(add(smul(1.0f - beta - gamma, nv(idx[3 * tri])),
                                               smul(beta, nv(idx[3 * tri + 1]))),
                                           smul(gamma, nv(idx[3 * tri + 2]))));
                    }
                    v3 d = cosine_dir(ng, ns, r.d, xi0, xi1);
                    T = v3{T.x * alb.x * scale, T.y * alb.y * scale, T.z * alb.z * scale};
                    if (T.x == 0.0f && T.y == 0.0f && T.z == 0.0f) break;
                    r = make_ray(xh, d, 1e-4f, FLT_MAX);
                }
            }
            if (sig) sig[p] = h;
        }
#pragma omp critical
        for (int L = 0; L < levels; ++L) counts[(size_t)L] += mine[(size_t)L];
    }
    if (level_rays)
        for (int L = 0; L < levels; ++L) level_rays[L] += counts[(size_t)L];
}

} // extern "C"

namespace {

// ---- wavefront direct lighting -------------------------------------------
//
// Lights: scene::find_light_geometry (libgi/scene.cpp:366-370) -- every
// emissive triangle in triangle order -- then point lights as the pointlight
// command appends them (driver/interaction.cpp:723-753).
struct olight {
    bool point;
    v3 a, b, c;      // triangle corners, or a = position
    v3 col;          // emission or point colour
    uint32_t tri;    // the emitting triangle
};

inline float luma3(v3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }   // libgi/color.h

struct olights {
    std::vector<olight> L;
    std::vector<float> f, cdf;
    float integral = 0;
    // scene::compute_light_distribution (libgi/scene.cpp:373-412) with
    // trianglelight::power / pointlight::power (scene.cpp:494-527) and
    // distribution_1d::build_cdf (libgi/discrete_distributions.cpp:25-38)
    void build() {
        const float pi = float(M_PI);
        for (const olight &l : L) {
            v3 p;
            if (l.point) {
                p = smul(4 * pi, l.col);
            } else {
                v3 e1 = sub(l.b, l.a), e2 = sub(l.c, l.a);
                v3 x = cross(e1, e2);
                p = mul(mul(mul(l.col, 0.5f), sqrtf(dot(x, x))), pi);
            }
            f.push_back(luma3(p));
        }
        size_t N = f.size();
        cdf.assign(N + 1, 0.0f);
        for (size_t i = 1; i < N + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1];
        integral = cdf[N];
        for (size_t i = 1; i < N + 1; ++i) cdf[i] = integral > 0 ? cdf[i] / integral : i / float(N);
    }
    // distribution_1d::sample_index (discrete_distributions.cpp:40-48)
    uint32_t sample_index(float xi, float &pdf) const {
        unsigned index = unsigned(std::lower_bound(cdf.begin(), cdf.end(), xi) - cdf.begin());
        index = index > 0 ? index - 1 : index;
        pdf = integral > 0.0f ? f[index] / integral : 0.0f;
        return index;
    }
};

struct direct_out {
    ray shadow;
    bool valid = false;   // shadow ray enabled (t_max > 0)
    float pdf = 0;
    v3 lightcol{0, 0, 0};
};

} // namespace

extern "C" {

// The wavefront direct-lighting algorithm of gi/direct.cpp:483-552 for the
// camera rays of `cam` (pixel centres), one pass per seed in `seeds`:
// closest hit of the camera ray, then by mode
//   0 uniform: sample_uniform_dir (rt/cpu/bounce.cpp:11-39), closest hit,
//              integrate_dir_sample (bounce.cpp:108-130);
//   1 cosine:  sample_cos_weighted_dir (bounce.cpp:41-72), the same;
//   2 light:   sample_light_dir (bounce.cpp:74-106) with trianglelight /
//              pointlight ::sample_Li (libgi/scene.cpp:500-559), any hit,
//              integrate_light_sample (bounce.cpp:132-151);
// accumulating (rgb, weight) per pixel into fb (4 floats per pixel) in the
// reference's order: the sample step's (emission, 1 -- or 0 in light mode)
// for camera hits on emitters, then the integration step's (radiance, 1).
// Random numbers per pixel i and seed s: PCG32 stream (s, i, 0); light mode
// draws the light index, then the two for the point on the light.
// Lambertian BRDF (lambertian_reflection::f, libgi/material.cpp:55-65).  The
// shading normal ns is the corners' normals vn interpolated at the hit
// (diff_geom::ns), or, with vn null, the geometric normal (the face normals
// the reference's loader generates for meshes without normals).  The sample
// steps flip the normals to the camera ray's side by ns (flip_normals_to_ray)
// and align their samples to the flipped geometric normal; the integration
// steps weigh by ns unflipped, as the reference does; a triangle light's
// normal at its sampled point is its corners' normals interpolated there
// (trianglelight::sample_Li), or its geometric normal.  points: n_points x 6
// floats (pos, col); mats: 8 floats per material; mat per triangle.
void orc_direct(const float *pos, const uint32_t *idx, uint32_t n_tris, const uint32_t *mat, const float *mats,
                const float *vn, const void *nodes, uint32_t n_nodes, uint32_t root, const uint32_t *index,
                const float *cam, int mode, const float *points, uint32_t n_points, const uint64_t *seeds,
                uint32_t n_seeds, float *fb, int threads) {
    scene_view s{pos, idx, n_tris};
    tree tr{static_cast<const node *>(nodes), n_nodes, root, index};
    pinhole ph(cam);
    const float pi = float(M_PI);
    const float one_over_pi = (1.0 / M_PI), one_over_2pi = (1.0 / (2 * M_PI));
    auto material = [&](uint32_t t) { return mats + 8 * (size_t)(mat ? mat[t] : 0u); };
    auto corner = [&](uint32_t t, int k) { return s.p(idx[3 * t + k]); };
    olights lights;
    for (uint32_t t = 0; t < n_tris; ++t) {
        const float *m = material(t);
        v3 e{m[4], m[5], m[6]};
        if (e.x != 0 || e.y != 0 || e.z != 0) lights.L.push_back({false, corner(t, 0), corner(t, 1), corner(t, 2), e, t});
    }
    for (uint32_t k = 0; k < n_points; ++k) {
        const float *p = points + 6 * k;
        lights.L.push_back({true, v3{p[0], p[1], p[2]}, v3{0, 0, 0}, v3{0, 0, 0}, v3{p[3], p[4], p[5]}, 0});
    }
    lights.build();
    const int64_t n_px = (int64_t)ph.w * ph.h;
#pragma omp parallel for schedule(dynamic, 256) num_threads(set_threads(threads))
    for (int64_t i = 0; i < n_px; ++i) {
        int x = (int)(i % ph.w), y = (int)(i / ph.w);
        float *px = fb + 4 * i;
        for (uint32_t sd = 0; sd < n_seeds; ++sd) {
            ray cr = ph.at(x, y, 0.0f, 0.0f);
            hit is = closest_hit(s, tr, cr, nullptr);
            direct_out o;
            o.pdf = mode == 0 ? one_over_2pi : (mode == 1 ? one_over_pi : 0.0f);
            v3 ng{0, 0, 0}, ns{0, 0, 0}, xh{0, 0, 0}, albedo{0, 0, 0};
            if (is.t != FLT_MAX) {
                v3 A = corner(is.ref, 0), B = corner(is.ref, 1), C = corner(is.ref, 2);
                const float *m = material(is.ref);
                albedo = v3{m[0], m[1], m[2]};
                v3 e{m[4], m[5], m[6]};
                xh = add(add(smul(1.0f - is.beta - is.gamma, A), smul(is.beta, B)), smul(is.gamma, C));
                ng = normalize(cross(sub(B, A), sub(C, A)));   // diff_geom::ng (libgi/rt.cpp:7-15)
                ns = vn ? shading_normal(vn, idx, is.ref, is.beta, is.gamma) : ng;
                if (dot(cr.d, ns) > 0) ng = v3{-ng.x, -ng.y, -ng.z};   // flip_normals_to_ray (util.h:226-233)
                if (e.x != 0 || e.y != 0 || e.z != 0) {
                    px[0] += e.x;
                    px[1] += e.y;
                    px[2] += e.z;
                    px[3] += mode == 2 ? 0.0f : 1.0f;
                } else {
                    pcg32 g(stream_of(seeds[sd], (uint64_t)i, 0));
                    if (mode == 0 || mode == 1) {
                        float xi0 = g.next_float(), xi1 = g.next_float();
                        v3 sdir;
                        if (mode == 0) {   // uniform_sample_hemisphere (libgi/sampling.h:19-26)
                            float z = xi0, r0 = 1 - z * z;
                            float r = sqrtf(r0 < 0 ? -r0 : r0);
                            float phi = 2 * pi * xi1;
                            sdir = v3{r * cosf(phi), r * sinf(phi), z};
                        } else {   // cosine_sample_hemisphere (sampling.h:32-36)
                            float r = sqrtf(xi0), theta = 2 * pi * xi1;
                            float dx = r * cosf(theta), dy = r * sinf(theta);
                            float z = 1 - dx * dx - dy * dy;
                            sdir = v3{dx, dy, z > 0 ? sqrtf(z) : 0.f};
                        }
                        // align (util.h:212-218)
                        float sg = copysignf(1.f, ng.z);
                        v3 w{sdir.x, sdir.y, sdir.z * sg}, hv{ng.x, ng.y, ng.z + sg};
                        float k = dot(w, hv) / (1.f + (ng.z < 0 ? -ng.z : ng.z));
                        v3 wi = sub(smul(k, hv), w);
                        if (mode == 1) {
                            float c = dot(wi, ng);
                            o.pdf *= c < 0.0f ? 0.0f : c;   // cdot
                        }
                        o.shadow = make_ray(xh, wi, 1e-4f, FLT_MAX);
                        o.valid = true;
                    } else {
                        float l_pdf;
                        uint32_t l = lights.sample_index(g.next_float(), l_pdf);
                        float xi0 = g.next_float(), xi1 = g.next_float();
                        const olight &L = lights.L[l];
                        v3 col = L.col, wi;
                        float a_pdf = 1.0f, tmax;
                        if (L.point) {   // pointlight::sample_Li (scene.cpp:500-508)
                            wi = sub(L.a, xh);
                            tmax = sqrtf(dot(wi, wi));
                            wi = v3{wi.x / tmax, wi.y / tmax, wi.z / tmax};
                            col = v3{col.x / (tmax * tmax), col.y / (tmax * tmax), col.z / (tmax * tmax)};
                        } else {   // trianglelight::sample_Li (scene.cpp:534-559)
                            float su0 = sqrtf(xi0);
                            float bx = 1.0f - su0, by = xi1 * su0;   // uniform_sample_triangle (sampling.h:43-46)
                            v3 target = add(add(smul(1.0f - bx - by, L.a), smul(bx, L.b)), smul(by, L.c));
                            v3 nl = vn ? shading_normal(vn, idx, L.tri, bx, by)
                                       : normalize(cross(sub(L.b, L.a), sub(L.c, L.a)));
                            wi = sub(target, xh);
                            v3 x2 = cross(sub(L.b, L.a), sub(L.c, L.a));
                            float area = 0.5f * sqrtf(dot(x2, x2));
                            tmax = sqrtf(dot(wi, wi));
                            wi = v3{wi.x / tmax, wi.y / tmax, wi.z / tmax};
                            float cos_l = dot(nl, v3{-wi.x, -wi.y, -wi.z});
                            if (cos_l <= 0.0f) {
                                col = v3{0, 0, 0};
                                a_pdf = 0.0f;
                            } else {
                                a_pdf = tmax * tmax / (cos_l * area);
                            }
                        }
                        if (col.x != 0 || col.y != 0 || col.z != 0) {
                            o.shadow = make_ray(xh, wi, 1e-4f, tmax - 1e-4f);   // ray::length_exclusive
                            o.valid = o.shadow.t_max > 0;
                            o.lightcol = col;
                        }
                        o.pdf = a_pdf * l_pdf;
                    }
                }
            }
            // integration
            v3 rad{0, 0, 0};
            if (is.t != FLT_MAX && o.valid) {
                v3 w_in = o.shadow.d;
                float c = dot(w_in, ns);
                float cosv = c < 0.0f ? 0.0f : c;
                v3 fb_ = dot(w_in, ns) > 0 ? smul(one_over_pi, albedo) : v3{0, 0, 0};   // lambertian f
                if (mode == 2) {
                    if (!any_hit(s, tr, o.shadow)) {
                        v3 lf = v3{o.lightcol.x * fb_.x, o.lightcol.y * fb_.y, o.lightcol.z * fb_.z};
                        v3 lc = mul(lf, cosv);
                        rad = v3{lc.x / o.pdf, lc.y / o.pdf, lc.z / o.pdf};
                    }
                } else {
                    hit lh = closest_hit(s, tr, o.shadow, nullptr);
                    if (lh.t != FLT_MAX) {
                        const float *lm = material(lh.ref);
                        v3 e{lm[4], lm[5], lm[6]};
                        if (e.x != 0 || e.y != 0 || e.z != 0) {
                            v3 ef = v3{e.x * fb_.x, e.y * fb_.y, e.z * fb_.z};
                            v3 ec = mul(ef, cosv);
                            rad = v3{ec.x / o.pdf, ec.y / o.pdf, ec.z / o.pdf};
                        }
                    }
                }
            }
            px[0] += rad.x;
            px[1] += rad.y;
            px[2] += rad.z;
            px[3] += 1.0f;
        }
    }
}

// Wall-clock seconds of one orc_trace_closest call (the CPU baseline leg).
double orc_time_closest(const float *pos, const uint32_t *idx, uint32_t n_tris, const void *nodes, uint32_t n_nodes,
                        uint32_t root, const uint32_t *index, const float *const *ray_planes, uint32_t n, uint32_t row,
                        float *t, float *u, float *v, uint32_t *id, int threads) {
    double t0 = omp_get_wtime();
    orc_trace_closest(pos, idx, n_tris, nodes, n_nodes, root, index, ray_planes, n, row, t, u, v, id, nullptr,
                      threads);
    return omp_get_wtime() - t0;
}

} // extern "C"

namespace {

// Uniform point of the ball of radius R about the origin from three PCG32
// numbers, as setup_ray_incoherent draws it (rt/cuda/kernels.cu:88-103):
// phi = 2 pi xi1, cos theta = 2 xi2 - 1, r = R cbrt(xi3).
v3 ball_offset(pcg32 &g, float R) {
    const float pi = float(M_PI);
    float xi1 = g.next_float(), xi2 = g.next_float(), xi3 = g.next_float();
    float phi = xi1 * 2 * pi;
    float theta = acosf(xi2 * 2 - 1);
    float r = R * cbrtf(xi3);
    return v3{r * sinf(theta) * cosf(phi), r * sinf(theta) * sinf(phi), r * cosf(theta)};
}

} // namespace

extern "C" {

// The incoherent benchmark rays of the cuda platform's camera step
// (setup_ray_incoherent, rt/cuda/kernels.cu:75-156): origin = s1 + a ball
// point of radius r1 * r_max (s1 itself for r1 = 0), direction = (s2 + a ball
// point of radius r2 * r_max) - origin, not normalized (kernels.cu:144 drops
// normalize's result), t in [0, FLT_MAX].  Ray i draws from the stream
// stream_of(seed, i, 0), origin first; the reference's cuRAND MTGP32 numbers
// are not reproducible here, so the geometry is what is pinned.
void orc_incoherent_rays(const float *s1, const float *s2, float r1, float r2, float r_max, uint64_t seed, uint32_t n,
                         float *const *out) {
    for (uint32_t i = 0; i < n; ++i) {
        pcg32 g(stream_of(seed, i, 0));
        v3 o{s1[0], s1[1], s1[2]}, t{s2[0], s2[1], s2[2]};
        if (r1 > 0) o = add(o, ball_offset(g, r1 * r_max));
        if (r2 > 0) t = add(t, ball_offset(g, r2 * r_max));
        const float row[8] = {o.x, o.y, o.z, 0.0f, t.x - o.x, t.y - o.y, t.z - o.z, FLT_MAX};
        for (int k = 0; k < 8; ++k) out[k][i] = row[k];
    }
}

} // extern "C"
