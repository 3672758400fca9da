// Device-side data layout of a flattened scene, shared by the host flattener
// and the gfx950 kernels.
//
// BVH2 node (64 B, four 16-B rows, the Aila-Laine arrangement also used by the
// reference's compact_bvh_node, rt/cuda/base.h:66-78):
//   row0 = (l.lo.x, l.hi.x, l.lo.y, l.hi.y)
//   row1 = (r.lo.x, r.hi.x, r.lo.y, r.hi.y)
//   row2 = (l.lo.z, l.hi.z, r.lo.z, r.hi.z)
//   row3 = (ref_l, ref_r, 0, 0)                as int32 bits
// A child reference >= 0 is the index of an inner node.  A leaf is encoded in
// its parent's reference as ~((slot << kLeafCountBits) | (count - 1)), so the
// kernel never fetches a leaf record: the triangle slots [slot, slot+count)
// come straight out of the reference.  (The reference instead stores the leaf
// as a node and flags it with a negative index.)
//
// BVH4 node (128 B, eight 16-B rows): the four child boxes as
//   row0 = lo.x[0..3]  row1 = hi.x[0..3]  row2 = lo.y[0..3]  row3 = hi.y[0..3]
//   row4 = lo.z[0..3]  row5 = hi.z[0..3]  row6 = ref[0..3]   row7 = unused
// An absent child has ref = kInvalidRef and an empty box (lo = +inf, hi = -inf).
//
// Triangle records, one per leaf slot (ESC may list a triangle in several
// leaves; every slot gets its own record), 48 B, 16-B aligned:
//   exact:  row0 = (A.xyz, tri_id bits)  row1 = (A-B).xyz  row2 = (A-C).xyz
//           -- holds exactly the differences the Shirley/Cramer test of
//              libgi/intersect.h:19-70 computes, so the kernel reproduces the
//              CPU test bit for bit.
//   woop:   row0..row2 = the three rows of the inverse of [B-A, C-A, N],
//           computed in FP64 and rounded, with w = A.x, A.y, A.z (the
//           triangle's local origin); tri ids live in a separate per-slot
//           array.
#pragma once

#include <cstdint>

namespace rt {

constexpr int kLeafCountBits = 5;                 // up to 32 slots per leaf
constexpr uint32_t kMaxLeafCount = 1u << kLeafCountBits;
constexpr uint32_t kMaxLeafSlot = (1u << (31 - kLeafCountBits)) - 1;
constexpr int32_t kInvalidRef = 0x7fffffff;       // also the traversal sentinel

inline constexpr int32_t make_leaf_ref(uint32_t slot, uint32_t count) {
    return ~(int32_t)((slot << kLeafCountBits) | (count - 1));
}
inline constexpr bool ref_is_leaf(int32_t ref) { return ref < 0; }
inline constexpr uint32_t leaf_slot(int32_t ref) { return ((uint32_t)~ref) >> kLeafCountBits; }
inline constexpr uint32_t leaf_count(int32_t ref) { return (((uint32_t)~ref) & (kMaxLeafCount - 1)) + 1; }

enum tri_format : int { TRI_EXACT = 0, TRI_WOOP = 1 };
enum bvh_format : int { BVH2 = 0, BVH4 = 1 };

struct alignas(16) node2 {
    float lxlo, lxhi, lylo, lyhi;
    float rxlo, rxhi, rylo, ryhi;
    float lzlo, lzhi, rzlo, rzhi;
    int32_t ref_l, ref_r, pad0, pad1;
};
static_assert(sizeof(node2) == 64, "node2 must be 64 B");

struct alignas(16) node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t ref[4];
    int32_t pad[4];
};
static_assert(sizeof(node4) == 128, "node4 must be 128 B");

struct alignas(16) tri_rec {
    float r0[4], r1[4], r2[4];
};
static_assert(sizeof(tri_rec) == 48, "tri_rec must be 48 B");

// Ray and hit buffers, structure of arrays.  32 B in and 16 B out per ray.
struct ray_soa {
    const float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
};
struct hit_soa {
    float *t, *u, *v;
    uint32_t *id;
};

// Everything a trace kernel needs to know about one device-resident scene.
struct scene_desc {
    const void *nodes;        // node2[] or node4[]
    const tri_rec *tris;      // one record per leaf slot
    const uint32_t *slot_ids; // slot -> triangle id (woop format; exact keeps it in row0.w)
    int32_t root_ref;         // inner node index, or a leaf reference for one-leaf trees
    int32_t bvh;              // bvh_format
    int32_t tri;              // tri_format
    int32_t depth;            // longest root-to-leaf path, in inner nodes
    int32_t nan_free;         // no node box bound is NaN (kernels may use the min/max box test)
    // Alpha-aware tracing (the reference's *-alpha tracers): when `alpha` is
    // set, a triangle hit counts only where the opacity of its material's
    // texture at the hit exceeds 0.5 (alpha_opaque in device/trace_core.h).
    // Per triangle two float4 (a.u, a.v, b.u, b.v | c.u, c.v, texture index
    // bits or ~0u, 0); per texture (first texel, w, h, 0); one byte per texel,
    // 1 where its alpha > 0.5.  Null when the scene has no textures.
    const void *alpha_tri;    // float4 pairs
    const void *alpha_tex;    // uint4
    const uint8_t *alpha_mask;
    int32_t alpha;
};

} // namespace rt
