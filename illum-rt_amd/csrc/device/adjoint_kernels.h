// Launcher of the closest-hit adjoint kernel (device/adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../core/layout.h"
#include "../core/vec.h"

namespace rt {
namespace dev {

struct adjoint_args {
    ray_soa rays;
    hit_soa hits;               // frozen forward hits: t (validity) and id
    const float4 *tri_geo;      // A + material, B, C per triangle
    const uint32_t *tri_vidx;   // vertex indices per triangle
    const float *grad_t, *grad_u, *grad_v;
    float *grad_vertices;       // (n_vert, 3), accumulated; may be null
    float *grad_origin;         // (n, 3); may be null
    float *grad_dir;            // (n, 3); may be null
    float *grad_corners;        // (n, 9): dA, dB, dC per ray; may be null
    uint32_t n_tris = 0;        // triangles of the scene (the sort key of a miss)
    uint32_t *keys = nullptr;   // set by the launcher
    uint32_t *count = nullptr;  // set by the launcher
    float *corner_scratch = nullptr;   // set by the launcher: 9 floats per ray when grad_corners is null
};

// Device bytes of the scratch the vertex accumulation sorts in, for n rays.
size_t adjoint_sort_bytes(uint32_t n);

// sort_scratch: adjoint_sort_bytes(n) bytes of device memory, needed when
// grad_vertices is set (the stream orders reuses).
hipError_t launch_adjoint_closest(const adjoint_args &a, uint32_t n, void *sort_scratch, hipStream_t s);

} // namespace dev
} // namespace rt
