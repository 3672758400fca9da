// Flattening of a host BVH into the device layout of core/layout.h.
//
// The reference uploads its node array as-is with a per-link leaf flag
// (compact_bvh_node_builder, rt/cuda/base.cpp:157-178) and reaches triangles
// through index -> triangle -> three vertex fetches.  Here leaves disappear
// into their parents' references, only inner nodes are stored (renumbered in
// the original pre-order, so traversal order is unchanged), and every leaf
// slot owns a 48-byte triangle record, so a triangle test is three 16-byte
// loads from one cache line pair.
#pragma once

#include "../core/layout.h"
#include "../core/records.h"
#include "bvh.h"
#include "scene.h"

#include <string>
#include <vector>

namespace rt {

struct flat_scene {
    int bvh_kind = BVH2;
    int tri_kind = TRI_EXACT;
    int32_t root_ref = kInvalidRef;   // kInvalidRef: no triangles
    uint32_t depth = 0;               // inner nodes on the longest root-leaf path
    std::vector<node2> nodes2;
    std::vector<node4> nodes4;
    std::vector<tri_rec> tris;        // one per leaf slot
    std::vector<uint32_t> slot_ids;   // slot -> triangle id
};

// Returns "" on success, else a message (leaf too large, too many slots).
std::string flatten(const scene &s, const bvh &b, int bvh_kind, int tri_kind, flat_scene &out);

} // namespace rt
