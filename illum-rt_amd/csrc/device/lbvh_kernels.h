// Device BVH construction (linear BVH) straight into the flattened layout of
// core/layout.h.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/layout.h"

#include <cstdint>

namespace rt {
namespace dev {

struct lbvh_in {
    const float *pos;        // 3 floats per vertex
    const uint32_t *vidx;    // 3 vertex indices per triangle
    uint32_t n_vert, n_tris;
    int max_leaf;            // leaf size bound (1..32)
    int tri;                 // tri_format of the records
};

struct lbvh_out {
    node2 *nodes;            // n_tris - 1 entries reserved (at most that many are written)
    tri_rec *tris;           // n_tris records, slot k = k-th triangle in Morton order
    uint32_t *slot_ids;      // n_tris
    int32_t *root_ref;       // device int32: root inner node 0, or the leaf reference of a one-leaf tree
    uint32_t *counts;        // device uint32[4]: emitted inner nodes, depth (inner nodes on the longest path),
                             // bad vertex indices, unused
};

// Scratch the build needs, in bytes, for n_tris triangles.
size_t lbvh_scratch_bytes(uint32_t n_tris);

// Builds the tree on `stream`; `scratch` holds lbvh_scratch_bytes(n_tris)
// bytes.  Results are valid once the stream reaches the end of the build.
hipError_t launch_lbvh(const lbvh_in &in, const lbvh_out &out, void *scratch, hipStream_t stream);

} // namespace dev
} // namespace rt
