// CPU binary BVH construction: object median, spatial median and the
// reference's plane-sweep SAH, with optional early split clipping (ESC).
//
// This is the builder the reference runs on the host for every platform
// (rt/cpu/bvh-ctor.h); its output -- node array, root and leaf-slot index --
// is what the GPU layout is flattened from and what the CPU tracer walks, so
// the restatement follows the reference's steps exactly, including its
// in-place partition, the degenerate-split fallback and the ESC threshold.
#pragma once

#include "../core/vec.h"
#include "scene.h"

#include <cstdint>
#include <string>
#include <vector>

namespace rt {

// rt/cpu/bvh.h:42-51: inner nodes hold two child indices (>= 0); a leaf holds
// link_l = -first_slot and link_r = -count.  56 bytes.
struct bbvh_node {
    aabb box_l, box_r;
    int32_t link_l, link_r;
    bool inner() const { return link_r >= 0; }
    int32_t tri_offset() const { return -link_l; }
    int32_t tri_count() const { return -link_r; }
};
static_assert(sizeof(bbvh_node) == 56, "bbvh_node must match the reference's 56-byte node");

// rt/cpu/bvh.h:58-63 (indexed layout): index[slot] = triangle id.
struct bvh {
    uint32_t root = 0;
    std::vector<bbvh_node> nodes;
    std::vector<uint32_t> index;
};

enum class split_kind { om, sm, sah };

struct bvh_params {
    split_kind split = split_kind::sah;
    int max_tris = 4;    // leaf size limit
    int planes = 16;     // SAH candidate planes
    bool esc = true;     // early split clipping
};

// Defaults of the reference's GPU platforms (rt/cuda/base.cpp:94-105): SAH,
// 16 planes, at most 4 triangles per leaf, ESC on.
inline bvh_params gpu_default_params() { return bvh_params{}; }
// Defaults of the reference's CPU binary_bvh_tracer (rt/cpu/bvh.h:77-81):
// object median, one triangle per leaf, one plane (used only by SAH).
inline bvh_params cpu_default_params() { return bvh_params{split_kind::om, 1, 1, false}; }

// Builds the BVH.  A scene without triangles yields an empty node array.
bvh build_bvh(const scene &s, const bvh_params &p);

// The ESC polygon splitter of rt/cpu/bvh-ctor.h:120-179, exposed for tests:
// fragment boxes of `poly` whose surface area is at most `threshold`.
std::vector<aabb> esc_split(const std::vector<vec3f> &poly, float threshold);

struct bvh_stats {
    uint32_t inner = 0, leaves = 0, slots = 0, max_leaf = 0, depth = 0;
};
bvh_stats compute_stats(const bvh &b);

// Checks structural invariants (links in range, every leaf slot in range,
// every triangle referenced, leaf sizes within max_tris, child boxes enclose
// the boxes of their subtrees' triangles).  Returns "" when all hold.
std::string validate_bvh(const scene &s, const bvh &b, int max_tris);

} // namespace rt
