// Device refit of a committed scene to new vertex positions: same triangles,
// same tree topology, new triangle records and node boxes.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/layout.h"

#include <cstdint>

namespace rt {
namespace dev {

struct refit_args {
    void *nodes;               // node2[] or node4[]
    int bvh;                   // bvh_format
    tri_rec *tris;             // one record per leaf slot
    const uint32_t *slot_ids;  // slot -> triangle id
    int tri;                   // tri_format
    uint32_t n_slots;
    float *tri_geo;            // 12 floats per triangle (A.xyz + material bits, B.xyz, C.xyz), or null
    uint32_t n_tris;
    const float *pos;          // 3 floats per vertex
    const uint32_t *vidx;      // 3 vertex indices per triangle
};

// Rewrites every triangle record (and the shading copy of the vertices) from
// `pos`.
hipError_t launch_refit_records(const refit_args &a, hipStream_t stream);

// Recomputes the child boxes of the `n` inner nodes listed at `list`; each
// listed node's inner children must already be final.
hipError_t launch_refit_level(const refit_args &a, const int32_t *list, uint32_t n, hipStream_t stream);

} // namespace dev
} // namespace rt
