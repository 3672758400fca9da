/*
 * illum_hiprt.h -- C ABI of the MI355X (gfx950) batch ray tracer.
 *
 * This is the drop-in boundary for the reference's batch ray-tracer path:
 * the C++ platform classes (wf::hip::platform and its tracers, mirroring
 * libgi/wavefront-rt.h) and the Python test/bench harness call only these
 * functions.  All entry points return 0 (HIPRT_OK) on success or a nonzero
 * status; the message of the last failure on the calling thread is available
 * from hiprt_last_error().  Nothing throws across this boundary.
 *
 * Conventions
 *   - Pointers named d_* are device pointers; all others are host pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Trace calls are asynchronous on that stream.
 *   - Rays are structure-of-arrays fp32: origin, t_min, direction, t_max
 *     (32 B per ray); hits are structure-of-arrays t, beta, gamma, triangle id
 *     (16 B per ray).  A miss is t = FLT_MAX, id = 0, as
 *     triangle_intersection::reset() (libgi/rt.h:57-67).
 */
#ifndef ILLUM_HIPRT_H
#define ILLUM_HIPRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HIPRT_OK = 0,
    HIPRT_ERR_ARG = 1,         /* invalid argument or handle */
    HIPRT_ERR_HIP = 2,         /* a HIP runtime call failed */
    HIPRT_ERR_UNSUPPORTED = 3, /* valid request this build cannot serve */
    HIPRT_ERR_IO = 4           /* file could not be read or written */
};

/* Message of the last failing call on this thread ("" if none).
 * Replaces CHECK_CUDA_ERROR's stderr+throw (rt/cuda/cuda-helpers.cpp:5-12). */
const char *hiprt_last_error(void);

/* Version string of this library. */
const char *hiprt_version(void);

/* ------------------------------------------------------------------ host rays
 * Camera rays of cam_ray (libgi/camera.cpp:44-61) for every pixel of a w x h
 * image, pixel-centre offset 0, t_min 1e-4, t_max FLT_MAX, written row-major
 * into eight host planes (ox, oy, oz, tmin, dx, dy, dz, tmax) of w*h floats.
 * cam = pos.xyz, dir.xyz, up.xyz, fovy (vertical, degrees), w, h. */
int hiprt_camera_rays_host(const float *cam, float *const *planes);

/* ---------------------------------------------------------------- host scene
 * A triangle scene in host memory: positions, index triples, per-triangle
 * material.  Stands for ::scene's vertices/triangles (libgi/scene.h) as far as
 * tracing needs them.  No GPU is touched. */
typedef struct hiprt_host_scene_s *hiprt_host_scene;

/* Synthetic workloads by number: 1 = 8,192 random triangles, 2 = interior hall
 * (~74k), 3 = dense hall + debris (262,144).  When cam is non-NULL it receives
 * 15 floats: pos.xyz, dir.xyz, up.xyz, fovy (degrees, vertical), w, h, and the
 * point light's position xyz. */
int hiprt_scene_generate(int config, hiprt_host_scene *out, float *cam);
/* n random triangles, centroids U[0,1]^3, vertex offsets N(0, sigma^2), PCG32 seed. */
int hiprt_scene_random(uint32_t n_tris, float sigma, uint64_t seed, hiprt_host_scene *out);
/* Copy of caller arrays: pos[3*n_vert], idx[3*n_tris], mat[n_tris] (mat may be NULL). */
int hiprt_scene_from_arrays(const float *pos, uint32_t n_vert, const uint32_t *idx, uint32_t n_tris,
                            const uint32_t *mat, hiprt_host_scene *out);
/* OBJ reader: the file as the reference's scene::add loads it through
 * assimp's OBJ importer (libgi/scene.cpp:292-356): triangles in the
 * importer's object/mesh order, materials "DefaultMaterial" then the
 * libraries' in definition order then undefined "usemtl" names in first-use
 * order (Kd 0.6 grey by default, albedo Kd^2.2), quads split from a concave
 * corner, winding swapped where it disagrees with the corners' normals.
 * Writer: positions, triangles and "usemtl m<id>" switches. */
int hiprt_scene_load_obj(const char *path, hiprt_host_scene *out);
int hiprt_scene_write_obj(hiprt_host_scene s, const char *path);
/* scene::add(path, name, trafo) (libgi/scene.cpp:292-356), the script
 * command "load <file> [name]" (driver/interaction.cpp:395-402): loads an OBJ
 * file as hiprt_scene_load_obj does and appends it to *scene (a new scene
 * when *scene is NULL).  name (NULL or "" for none) prefixes its material
 * names as "<name>/<material>"; trafo is a column-major 4x4 model matrix
 * (16 floats; NULL = identity) applied to positions, with its inverse
 * transpose applied to normals, before the winding check
 * (mesh_load_process_node, libgi/scene.cpp:222-289).  Material ids of the
 * new triangles follow the scene's existing materials. */
int hiprt_scene_add_obj(hiprt_host_scene *scene, const char *path, const char *name, const float *trafo);
int hiprt_scene_info(hiprt_host_scene s, uint32_t *n_vert, uint32_t *n_tris);
/* Any output pointer may be NULL. */
int hiprt_scene_copy(hiprt_host_scene s, float *pos, uint32_t *idx, uint32_t *mat);
/* Material table: 8 floats per material (albedo rgb, 0, emission rgb, 0), the
 * layout hiprt_scene_build takes; *n_materials = count (table may be NULL). */
int hiprt_scene_materials(hiprt_host_scene s, float *table, uint32_t *n_materials);
/* Name of material `material` (material::name, which "material select"
 * looks up, driver/interaction.cpp:624-635), NUL-terminated into name[size];
 * "" for a scene whose materials have no names.  An OBJ file's materials are
 * "DefaultMaterial", then the libraries' materials in definition order, then
 * names "usemtl" uses without a definition. */
int hiprt_scene_material_name(hiprt_host_scene s, uint32_t material, char *name, uint32_t size);
/* Sets material `material`'s albedo and/or emission (3 floats each; NULL
 * keeps the value), as "material albedo" / "material emissive" do
 * (driver/interaction.cpp:642-651).  Takes effect at the next commit. */
int hiprt_scene_set_material(hiprt_host_scene s, uint32_t material, const float *albedo, const float *emissive);
/* Per-vertex texture coordinates, n_vert (u, v) pairs: vertex::tc of
 * libgi/rt.h, which the reference fills from assimp (libgi/scene.cpp:253-262).
 * n_vert must be the scene's vertex count.  Only alpha-aware tracing reads
 * them.  Returns HIPRT_ERR_ARG on a count mismatch. */
int hiprt_scene_set_texcoords(hiprt_host_scene s, const float *tc, uint32_t n_vert);
/* The per-vertex texture coordinates (2 floats per vertex; zeros for a scene
 * without them); *n_vert = the vertex count; tc may be NULL.  An OBJ file's
 * "vt" coordinates arrive here: every distinct (position, vt) corner pair is
 * a vertex of its own, after the file's "v" vertices. */
int hiprt_scene_get_texcoords(hiprt_host_scene s, float *tc, uint32_t *n_vert);
/* Replaces the vertex positions (vertex::pos of libgi/rt.h), n_vert (x, y,
 * z) triples; n_vert must be the scene's vertex count.  Read at the next
 * build or commit; a committed scene moves only through hiprt_scene_refit.
 * The platform's scene steps "rotate scene" / "drop scene view"
 * (rt/cuda/preprocessing.cpp:41-72) edit positions this way. */
int hiprt_scene_set_positions(hiprt_host_scene s, const float *pos, uint32_t n_vert);
/* Per-vertex shading normals, n_vert (x, y, z) triples: vertex::norm of
 * libgi/rt.h, interpolated at a hit into diff_geom::ns (libgi/rt.cpp:7-15).
 * The wavefront steps flip normals to the ray's side by it
 * (flip_normals_to_ray, libgi/util.h:226-232), the Lambertian integration
 * weighs by it, and a triangle light's sample point takes its interpolated
 * normal (trianglelight::sample_Li, libgi/scene.cpp:534-559).  n_vert must be
 * the scene's vertex count, or 0 to remove them (the geometric normal then
 * shades).  Read at commit; a refit keeps the committed normals. */
int hiprt_scene_set_normals(hiprt_host_scene s, const float *normals, uint32_t n_vert);
/* The shading normals: *n_vert = the vertex count, or 0 for a scene without
 * them; normals (3 floats per vertex) may be NULL.  An OBJ file's "vn"
 * normals arrive here: every distinct (v, vt, vn) corner is a vertex of its
 * own, and a corner without "vn" in a file with normals gets its face's. */
int hiprt_scene_get_normals(hiprt_host_scene s, float *normals, uint32_t *n_vert);
/* Attaches an RGBA texture (w*h texels of 4 floats, texel (x, y) at
 * 4*(y*w + x)) to material `material` as its albedo texture, the
 * material::albedo_tex that libgi/scene.cpp:340-349 loads; rgba NULL detaches
 * it.  Alpha-aware tracers test its alpha channel (texture2d::sample's
 * nearest texel with wrap, libgi/scene.h:28-36).  Returns HIPRT_ERR_ARG for
 * a material id past the scene's table or an empty texture. */
int hiprt_scene_set_texture(hiprt_host_scene s, uint32_t material, const float *rgba, uint32_t w, uint32_t h);
void hiprt_scene_free(hiprt_host_scene s);

/* ------------------------------------------------------------------ host BVH
 * The reference's CPU BVH builders (rt/cpu/bvh-ctor.h): object median (0),
 * spatial median (1) or plane-sweep SAH (2), with optional early split
 * clipping.  Defaults of the reference's GPU platforms: {2, 4, 16, 1}
 * (rt/cuda/base.cpp:94-105). */
typedef struct {
    int32_t split;     /* 0 om, 1 sm, 2 sah */
    int32_t max_tris;  /* triangles per leaf */
    int32_t planes;    /* SAH candidate planes */
    int32_t esc;       /* early split clipping on/off */
} hiprt_bvh_params;

typedef struct hiprt_bvh_s *hiprt_bvh;

int hiprt_bvh_build(hiprt_host_scene s, const hiprt_bvh_params *params, hiprt_bvh *out);
/* n_nodes: entries of the reference node array; n_index: leaf slots;
 * depth: inner nodes on the longest root-to-leaf path. */
int hiprt_bvh_info(hiprt_bvh b, uint32_t *n_nodes, uint32_t *n_index, uint32_t *root, uint32_t *depth);
/* nodes: n_nodes records of 56 B in the reference's bbvh_node layout
 * (rt/cpu/bvh.h:42-52: box_l.min/max, box_r.min/max, link_l, link_r);
 * index: n_index uint32.  Either pointer may be NULL. */
int hiprt_bvh_copy(hiprt_bvh b, void *nodes, uint32_t *index);
/* Structural check of the tree against the scene; 0 if every invariant holds,
 * else HIPRT_ERR_ARG with the broken invariant in hiprt_last_error(). */
int hiprt_bvh_validate(hiprt_host_scene s, hiprt_bvh b, int32_t max_tris);
void hiprt_bvh_free(hiprt_bvh b);

/* ------------------------------------------------------------------- context */
typedef struct hiprt_context_s *hiprt_context;

/* Selects HIP device `device` (cudaSetDevice(0) of rt/cuda/platform.cpp:29). */
int hiprt_create(int32_t device, hiprt_context *out);
int hiprt_destroy(hiprt_context ctx);

/* ------------------------------------------------------------ device scenes
 * Layouts: bvh_format 0 = binary (64-B Aila-Laine nodes), 1 = collapsed 4-wide
 * (128-B nodes); tri_format 0 = exact (A, A-B, A-C records: bit-identical to
 * the CPU Shirley test), 1 = Woop (rows of the inverse edge matrix about the
 * corner A as a local origin).  48-B triangle records. */
enum { HIPRT_BVH2 = 0, HIPRT_BVH4 = 1 };
enum { HIPRT_TRI_EXACT = 0, HIPRT_TRI_WOOP = 1 };

typedef struct hiprt_scene_s *hiprt_scene;

/* Flattens (scene, bvh) on the host and uploads it: the work of
 * batch_rt::build + compact_bvh_node_builder (rt/cuda/base.cpp:88-117,157-178). */
int hiprt_scene_commit(hiprt_context ctx, hiprt_host_scene s, hiprt_bvh b, int32_t bvh_format,
                       int32_t tri_format, hiprt_scene *out);
/* Builds a linear BVH on the device (Morton codes, radix sort, binary radix
 * tree, bottom-up boxes; leaves of at most max_leaf triangles) from
 * triangles already in device memory -- positions (3 floats per vertex) and
 * vertex indices (3 per triangle) -- and commits it with exact or Woop
 * records, shading data (optional per-triangle material ids on the device,
 * host material table of 8 floats: albedo rgb, 0, emission rgb, 0) and vertex
 * indices, so tracing, path tracing, the adjoint and refit all work on it.
 * The device counterpart of the host build + commit that every commit runs
 * (rt/cuda/base.cpp:88-117).  This tree is not the reference's SAH tree:
 * hits agree with brute force up to equal-t ties and grazing hits.
 * *device_ms (optional) receives the build kernels' time on `stream`; the
 * call returns after the scene is ready (it copies the flattened tree back
 * to the host once to check it). */
int hiprt_scene_build(hiprt_context ctx, const float *d_positions, uint32_t n_vert, const uint32_t *d_indices,
                      uint32_t n_tris, const uint32_t *d_materials, const float *materials, uint32_t n_materials,
                      int32_t max_leaf, int32_t tri_format, hiprt_scene *out, float *device_ms, void *stream);
/* Uploads an already flattened scene (host arrays): n_nodes nodes of 64 B
 * (bvh_format 0) or 128 B (1); n_slots triangle records of 48 B; slot_ids[n_slots]
 * triangle id per slot; root_ref = index of the root inner node or the leaf
 * reference of a one-leaf tree.  The arrays are checked on the host (child
 * indices increase along every edge, leaf slots in range) before upload. */
int hiprt_scene_upload(hiprt_context ctx, const void *nodes, uint32_t n_nodes, int32_t bvh_format,
                       const void *tri_records, const uint32_t *slot_ids, uint32_t n_slots, int32_t tri_format,
                       int32_t root_ref, hiprt_scene *out);
/* Host copy of the flattened arrays (sizes from hiprt_scene_layout). */
int hiprt_scene_layout(hiprt_scene sc, uint32_t *n_nodes, uint32_t *n_slots, int32_t *root_ref, uint32_t *depth,
                       int32_t *bvh_format, int32_t *tri_format);
/* Device arrays of the scene (sizes from hiprt_scene_layout), after waiting
 * for the device: reflects refits. */
int hiprt_scene_download(hiprt_scene sc, void *nodes, void *tri_records, uint32_t *slot_ids);
/* Refit to new vertex positions (d_positions: 3 floats per vertex, as many
 * vertices as the committed host scene): same triangles and tree topology,
 * triangle records rebuilt, node boxes recomputed bottom-up on the device,
 * one launch per tree level.  Leaf boxes bound whole triangles.  The device
 * counterpart of rebuilding on every commit (rt/cuda/base.cpp:88-117) when
 * only vertices move -- scene-view transforms (rt/cuda/preprocessing.cpp:41-72)
 * or gradient steps on vertices.  Needs a scene from hiprt_scene_commit;
 * stream-ordered, no host synchronisation. */
int hiprt_scene_refit(hiprt_scene sc, const float *d_positions, void *stream);
int hiprt_scene_destroy(hiprt_scene sc);
/* Most traversal-stack entries any ray can need in this tree (the persistent
 * tracers need it to fit their stack configuration, the plain one 63). */
int hiprt_scene_stack_need(hiprt_scene sc, uint32_t *stack_need);

/* --------------------------------------------------------------------- trace */
typedef struct {
    const float *d_ox, *d_oy, *d_oz, *d_tmin, *d_dx, *d_dy, *d_dz, *d_tmax;
} hiprt_rays;

typedef struct {
    float *d_t, *d_u, *d_v;
    uint32_t *d_id;
} hiprt_hits;

/* Tracer variants, named as the reference registers its CUDA tracers
 * (rt/cuda/platform.cpp:34-56).  hiprt_tracer_id maps a name to an id, -1 if
 * unknown; "default" and "find closest hits" are "while-while", the links of
 * rt/cuda/platform.cpp:55-56.  The reference's other loop structures of the
 * same traversal name the kernel that runs it here:
 * "simple" and "if-if" -> while-while, "persistent-if-if" ->
 * persistent-while-while, "speculative-while-while" ->
 * persistent-speculative-while-while (each also with "-alpha");
 * hiprt_tracer_name returns the canonical name. */
enum {
    HIPRT_TRACER_WHILE_WHILE = 0,                        /* plain: one ray per lane, grid = batch */
    HIPRT_TRACER_PERSISTENT_WHILE_WHILE = 1,             /* persistent waves, one atomic per fetch */
    HIPRT_TRACER_PERSISTENT_SPECULATIVE_WHILE_WHILE = 2, /* + leaf postponement */
    HIPRT_TRACER_DYNAMIC_WHILE_WHILE = 3,                /* + leaf postponement + dynamic refill;
                                                            2 and 3 may return either triangle
                                                            of an equal-t tie, and which one
                                                            can vary between runs */
    HIPRT_TRACER_COUNT = 4,
    /* OR'ed into a tracer id: its alpha-aware variant, the reference's
     * "<tracer>-alpha" (rt/cuda/platform.cpp:107-122; CPU semantics of
     * rt/cpu/bvh2.cpp:109-119,171-176): a triangle hit counts only where the
     * alpha of its material's albedo texture at the hit's texture coordinate
     * exceeds 0.5 (materials without a texture are opaque).  Scenes committed
     * without textures trace as with the plain id. */
    HIPRT_TRACER_ALPHA = 0x100
};
/* Names map to ids and back, "<name>-alpha" to id | HIPRT_TRACER_ALPHA. */
int hiprt_tracer_id(const char *name);
const char *hiprt_tracer_name(int tracer);

/* Tunables for the persistent tracers (zero = default). */
typedef struct {
    int32_t refill_threshold; /* refill when fewer lanes of 64 are active (default 40) */
    int32_t blocks;           /* grid size (default: occupancy x CU count) */
    uint64_t *d_fetch_stats;  /* non-NULL: run the counting variant, which adds the inner
                                 nodes it fetches to [0] and the triangle records it
                                 tests to [1] (device uint64[2]) */
    int32_t stack_config;     /* traversal stack per lane (VGPR + LDS entries, both tree
                                 widths): 0 = 4 + 28, 1 = 8 + 16, 2 = 2 + 20, 3 = 0 + 24,
                                 4 = 0 + 20 with registers capped for 8 waves per SIMD
                                 (binary trees, single scene); deeper traversals spill
                                 to a global column per lane */
    int32_t work_chunks;      /* 1 = one work counter; 8 (and 0, the default) = one
                                 contiguous chunk of the rays per XCD, blocks start on
                                 their own XCD's chunk (fetch atomics spread over 8
                                 counters: +57% on short any-hit rays) */
    int32_t spill_limit;      /* > 0: at most this many spill-column entries per lane
                                 (bounds the spill memory); 0 = what the tree needs.  A
                                 traversal that outgrows its stack never writes out of
                                 bounds: its result is marked lost -- t = NaN, id
                                 0xFFFFFFFF, occlusion byte 0xFF -- for that ray alone */
    int32_t park_leaves;      /* speculative / dynamic closest hit: leaves a lane parks before
                                 it stops descending, 1..4 (0 = 4), the postponement
                                 threshold */
} hiprt_tracer_opts;

/* compute_closest_hit of batch_ray_tracer (libgi/wavefront-rt.h:145-149,
 * rt/cpu/wavefront.cpp:53-59): hits[i] = closest hit of ray i. */
int hiprt_trace_closest(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, const hiprt_hits *hits, int32_t tracer,
                        const hiprt_tracer_opts *opts, void *stream);
/* compute_any_hit (rt/cpu/wavefront.cpp:60-70): d_occluded[i] = 1 if any
 * triangle is hit in (t_min, t_max), else 0.  When hits is non-NULL its
 * arrays are also written the way the reference's CPU adapter does:
 * t = ref = 0xFFFFFFFF (as float 4294967295) if occluded, reset() otherwise. */
int hiprt_trace_any(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, uint8_t *d_occluded, const hiprt_hits *hits,
                    int32_t tracer, const hiprt_tracer_opts *opts, void *stream);
/* hiprt_trace_closest without any device allocation by the library.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  The persistent tracers' work counters and their
 * traversal-stack spill column live in a static per-device scratch of
 * 2 KiB + 16 MiB (held until the process ends, its address resolved by
 * hiprt_create, which must have run on the device); the caller passes rays
 * and hits as for hiprt_trace_closest and gets the same results.  A launch
 * whose spill column (grid blocks x 256 lanes x spill entries per lane x 4 B,
 * entries = the tree's stack need minus the stack configuration's VGPR + LDS
 * entries) exceeds 16 MiB is refused with HIPRT_ERR_UNSUPPORTED before any
 * launch.  Every call on one device shares that scratch: calls that overlap
 * in time on one device (concurrent streams, threads) race and must be
 * serialized by the caller; consecutive calls on one stream are fine (the
 * counters are cleared by a kernel at each launch). */
int hiprt_trace_closest_noalloc(hiprt_scene sc, const hiprt_rays *rays, uint32_t n, const hiprt_hits *hits,
                                int32_t tracer, const hiprt_tracer_opts *opts, void *stream);

/* ------------------------------------------------------------ batched trace
 * Many scenes, each with its own ray batch, traced by ONE persistent launch
 * through a descriptor table: the batches form one index space from which
 * every wavefront fetches rays, and each lane binds the scene of the ray it
 * fetched.  Entries whose scenes differ in BVH or triangle format are traced
 * by one launch per format group.  closest hit: hits required; any hit:
 * d_occluded and/or hits.  No reference counterpart (the reference traces one
 * scene per launch, rt/cuda/tracers.cu). */
typedef struct {
    hiprt_scene scene;
    hiprt_rays rays;
    hiprt_hits hits;
    uint8_t *d_occluded;
    uint32_t n;
} hiprt_batch_entry;

int hiprt_trace_batch(const hiprt_batch_entry *entries, uint32_t n_entries, int32_t any, int32_t tracer,
                      const hiprt_tracer_opts *opts, void *stream);

/* --------------------------------------------------------------- autotuner
 * Times candidate configurations on a sample of a scene's rays -- tracer
 * variant (persistent / speculative / dynamic), refill threshold, stack
 * configuration, work chunks and BVH layout (binary vs collapsed 4-wide), on
 * the occupancy-derived grid, then the parked-leaf limit of the winner -- and
 * returns the fastest.  A sample of fewer
 * than 2^20 rays is tiled up to that size before timing: a batch smaller than
 * the lanes a persistent grid holds (~300k) measures tail latency and ranks
 * configurations wrongly.  Every candidate's hits on the sample are compared
 * with the plain kernel's; a candidate that differs other than at equal t is
 * rejected.  Results never depend on the choice.  The scene is committed in
 * the chosen layout into *tuned (caller destroys it). */
typedef struct {
    int32_t tracer;
    int32_t refill_threshold;
    int32_t blocks;
    int32_t bvh_format;
    float mrays_per_s;     /* measured on the sample */
    int32_t candidates;    /* configurations timed */
    int32_t stack_config;
    int32_t work_chunks;
    int32_t park_leaves;   /* 1..4, timed for the speculative and dynamic tracers */
} hiprt_tune_result;

int hiprt_autotune(hiprt_context ctx, hiprt_host_scene s, hiprt_bvh b, const hiprt_rays *sample, uint32_t n,
                   int32_t any, hiprt_tune_result *result, hiprt_scene *tuned);

/* ----------------------------------------------------------- wavefront steps
 * GPU forms of the reference's wavefront steps around the trace:
 *   "sample camera rays"        (rt/cuda/rni.cpp:31-101, cam_ray semantics)
 *   "sample cos weighted dir"   (rt/cpu/bounce.cpp:41-72)
 * Output rays go to writable device planes. */
typedef struct {
    float *d_ox, *d_oy, *d_oz, *d_tmin, *d_dx, *d_dy, *d_dz, *d_tmax;
} hiprt_ray_buffers;

/* w*h camera rays (cam as for hiprt_camera_rays_host); bit-identical to it. */
int hiprt_camera_rays(const float *cam, const hiprt_ray_buffers *out, void *stream);
/* The camera step's incoherent benchmark rays under "incoherence r1 r2"
 * (launch_setup_ray_incoherent, rt/cuda/kernels.cu:75-165): n rays from
 * sphere1 (3 floats) moved by a uniform point of the ball of radius
 * r1 * r_max, towards sphere2 moved by a uniform point of the ball of radius
 * r2 * r_max; the direction is the unnormalized difference, t in
 * [0, FLT_MAX].  Ray i draws from the PCG32 stream of (seed, i) (the
 * reference uses cuRAND's MTGP32).  r1, r2, r_max >= 0. */
int hiprt_incoherent_rays(const float *sphere1, const float *sphere2, float r1, float r2, float r_max, uint64_t seed,
                          const hiprt_ray_buffers *out, uint32_t n, void *stream);

/* Stable radix sort of ray keys: the sort that orders the path tracer's
 * bounce queues (8 key bits per pass).  No counterpart in the reference; it
 * is for callers that order their own ray batches before tracing them (rays
 * keyed by direction and origin cell trace faster in key order) and for
 * tests of the queue order.  Sorts the n keys in d_keys (device memory; used
 * as scratch, contents undefined afterwards) by their low `bits` bits
 * (1..32): d_perm[j] receives the index of the key at sorted position j,
 * equal keys in index order.  d_workspace: hiprt_sort_keys_workspace(n)
 * bytes of device memory, 256-byte aligned.  Asynchronous on `stream`. */
size_t hiprt_sort_keys_workspace(uint32_t n);
int hiprt_sort_keys(uint32_t *d_keys, uint32_t *d_perm, uint32_t n, int bits, void *d_workspace, void *stream);
/* out[i] = cosine-weighted bounce ray off hit i of ray i (disabled ray
 * o = 0, d = (1,1,1), t_max = -FLT_MAX for misses).  PCG32 stream per
 * (seed, i, level).  Needs a scene committed from a host scene. */
int hiprt_bounce_cosine(hiprt_scene sc, const hiprt_rays *rays, const hiprt_hits *hits, uint32_t n, uint64_t seed,
                        int32_t level, const hiprt_ray_buffers *out, void *stream);

/* framebuffer[i] += (albedo of hit i's material, 1), (0,0,0,1) on a miss:
 * "add hitpoint albedo" (rt/cpu/wavefront.cpp:105-119).  d_framebuffer holds
 * 4 floats per ray (rgb, weight). */
int hiprt_add_hitpoint_albedo(hiprt_scene sc, const hiprt_hits *hits, uint32_t n, float *d_framebuffer, void *stream);
/* Shadow ray from hit i towards the point light: origin = hit point, unit
 * direction, t_max = distance - 1e-4 (ray::length_exclusive, libgi/rt.h:30);
 * the disabled ray for misses.  The point-light case of "sample light dir"
 * (rt/cpu/bounce.cpp:74-106). */
int hiprt_shadow_rays(hiprt_scene sc, const hiprt_rays *rays, const hiprt_hits *hits, uint32_t n, const float *light,
                      const hiprt_ray_buffers *out, void *stream);
/* framebuffer[i] += the direct light of one point light of colour col at hit
 * i when its shadow ray is valid and unoccluded, weight += 1: "integrate
 * light sample" (rt/cpu/bounce.cpp:132-151) with pointlight::sample_Li
 * (libgi/scene.cpp:500-508): col / dist^2 * albedo / pi * max(0, w_in . n),
 * n the hit triangle's (unflipped) normal, zero where w_in is behind it. */
int hiprt_integrate_light(hiprt_scene sc, const hiprt_rays *cam_rays, const hiprt_hits *hits,
                          const hiprt_rays *shadow_rays, const uint8_t *d_occluded, uint32_t n, const float *light,
                          const float *col, float *d_framebuffer, void *stream);

/* ------------------------------------------------- direct-lighting steps
 * The steps of the reference's wavefront direct lighting (gi/direct.cpp:
 * 483-552, rt/cpu/bounce.cpp), selected by its "is" modes.  Materials are
 * Lambertian; a hit's shading normal is its triangle's geometric normal.
 *
 * Light list (for HIPRT_DIRECT_LIGHT): every emissive triangle of the scene
 * in triangle order (scene::find_light_geometry, libgi/scene.cpp:366-370),
 * then the given point lights (6 floats each: pos xyz, col rgb), sampled in
 * proportion to the luma of their power (scene::compute_light_distribution,
 * libgi/scene.cpp:373-412; distribution_1d::sample_index).  *n_lights (may
 * be NULL) receives the count.  Synchronous (reads the committed triangles
 * back once). */
int hiprt_scene_set_lights(hiprt_scene sc, const float *point_lights, uint32_t n_points, uint32_t *n_lights);

enum { HIPRT_DIRECT_UNIFORM = 0, HIPRT_DIRECT_COSINE = 1, HIPRT_DIRECT_LIGHT = 2 };

/* Sample step for camera ray i and its closest hit (PCG32 stream (seed, i, 0)):
 *   UNIFORM  sample_uniform_dir (rt/cpu/bounce.cpp:11-39): a uniform
 *            direction about the hit's normal flipped to the camera side,
 *            pdf = 1/(2 pi);
 *   COSINE   sample_cos_weighted_dir (bounce.cpp:41-72): cosine-weighted,
 *            pdf = cos / pi;
 *   LIGHT    sample_light_dir (bounce.cpp:74-106): a light by power, then a
 *            point on it (trianglelight / pointlight ::sample_Li); the shadow
 *            ray ends 1e-4 before it; pdf = area pdf x light pdf; the light's
 *            colour into d_light_col (3 floats per ray).  Lights facing away
 *            give the disabled ray.
 * out: the sampled rays (the disabled ray o = 0, d = (1,1,1), t_max = -FLT_MAX
 * for misses and emitters); d_pdf: the pdf per ray.  A camera hit on an
 * emitter adds (emission, 1) to d_framebuffer (float4 rgb + weight per ray),
 * (emission, 0) in LIGHT mode, as the reference's steps do. */
int hiprt_sample_direct(hiprt_scene sc, int32_t mode, const hiprt_rays *cam_rays, const hiprt_hits *cam_hits,
                        uint32_t n, uint64_t seed, const hiprt_ray_buffers *out, float *d_pdf, float *d_light_col,
                        float *d_framebuffer, void *stream);
/* integrate_dir_sample (bounce.cpp:108-130): framebuffer[i] += (emission of
 * the surface shadow ray i hit x albedo/pi x max(0, w_in . n) / pdf, 1), zero
 * where w_in is behind the camera hit's normal or nothing emits there. */
int hiprt_integrate_dir_sample(hiprt_scene sc, const hiprt_rays *cam_rays, const hiprt_hits *cam_hits,
                               const hiprt_rays *shadow_rays, const hiprt_hits *shadow_hits, const float *d_pdf,
                               uint32_t n, float *d_framebuffer, void *stream);
/* integrate_light_sample (bounce.cpp:132-151): framebuffer[i] += (light
 * colour x albedo/pi x max(0, w_in . n) / pdf, 1) where shadow ray i is valid
 * (t_max > 0) and unoccluded (d_occluded from an any-hit trace). */
int hiprt_integrate_light_sample(hiprt_scene sc, const hiprt_rays *cam_rays, const hiprt_hits *cam_hits,
                                 const hiprt_rays *shadow_rays, const uint8_t *d_occluded, const float *d_pdf,
                                 const float *d_light_col, uint32_t n, float *d_framebuffer, void *stream);

/* ------------------------------------------------------- wavefront path tracer
 * A diffuse path tracer over one committed scene and camera, with the control
 * flow of simple_pt::path (gi/pt.cpp:72-115): camera rays (jittered within
 * the pixel per sample), closest hit; a
 * miss ends the path, an emitter adds throughput x emission and ends it; else
 * a cosine-weighted bounce (throughput times albedo; a zero throughput ends
 * the path) and Russian roulette with a fixed survival probability from
 * level rr_start on.  Two schedules with
 * identical per-path results:
 *   batched = 0: one launch per step per level over all w*h slots, dead paths
 *                kept as disabled rays (the reference's wavefront form);
 *   batched = 1: live paths compacted into a queue after every level, traced
 *                by a persistent kernel that reads the queue length on the
 *                device, the whole frame captured once in a hipGraph.
 * In the batched schedule the queue of every bounce is, by default
 * (ray_order = 0), sorted by direction octant and origin cell (a 16^3 grid
 * over the scene box) so that wavefronts trace neighbouring rays; 1 keeps
 * the order the shading step appends them in.  Paths keep their own random
 * streams, so the image does not depend on the order.  samples > 1 traces
 * that many samples per pixel in one frame graph -- render call k traces the
 * samples of frames k*samples .. k*samples + samples - 1, with their own
 * random streams -- so every launch holds samples times more rays; radiance
 * is summed per sample and added to the image in sample order. */
typedef struct {
    int32_t max_depth;   /* last trace level; 4 = primary + 4 bounces */
    int32_t rr_start;    /* first level whose ray is subject to roulette (3) */
    float rr_survive;    /* survival probability (0.8) */
    uint64_t seed;
    int32_t tracer;      /* tracer id, may carry HIPRT_TRACER_ALPHA; -1 = while-while (plain) / default (batched) */
    int32_t batched;
    int32_t stack_config;   /* persistent tracer options (batched), as in hiprt_tracer_opts; */
    int32_t work_chunks;    /* 0 in refill_threshold / blocks selects the default */
    int32_t refill_threshold;
    int32_t blocks;
    int32_t ray_order;      /* batched: 0 = sorted bounce queues, 1 = shading order */
    int32_t samples;        /* batched: samples per pixel traced together (0 or 1: one) */
    int32_t emitters_reflect; /* 0 (default): a path ends at the first emitter it meets, after
                                 adding throughput x emission (gi/pt.cpp:86-89); 1: emitters
                                 also reflect and the path goes on */
    int32_t park_leaves;    /* batched: parked leaves per lane (1..4, 0 = 4) */
    int32_t pixel_centre;   /* 0 (default): each sample's camera ray passes through a random point
                               of its pixel -- offsets in [-0.5, 0.5)^2 from the path's own PCG32
                               stream, as simple_pt::sample_pixel jitters (gi/pt.cpp:36);
                               1: through the pixel centre */
    int32_t regen_pool;     /* batched: 0 (default) traces the samples of a render call in flight
                               together, level by level; k > 0 runs path regeneration over a
                               fixed pool of k slots per pixel (at most `samples`): each slot
                               renders its pixel's samples one after another, and a path that
                               ends is replaced in the next iteration by the camera ray of its
                               slot's next sample, so every trace launch holds a full pool.
                               With k = 1 the image equals the in-flight schedule's bit for bit
                               (a pixel's samples end, and add, in sample order); with k > 1
                               each slot sums its own samples and the slots are added at the
                               end, which reassociates the per-pixel sum.  render() then
                               waits on the device: it reads the pool's queue length every
                               few iterations to know when the frame is done. */
    int32_t refill_coherent; /* dynamic tracer with samples in flight (samples > 1, regen_pool 0):
                                refill threshold of the camera level and the first bounce, whose
                                rays of one pixel share wavefronts (0 = 16); refill_threshold
                                applies to the later levels */
} hiprt_pt_params;

typedef struct hiprt_pt_s *hiprt_pt;

int hiprt_pt_create(hiprt_scene sc, const float *cam, const hiprt_pt_params *params, hiprt_pt *out);
/* Renders one frame (one sample per pixel) into the accumulator, async. */
int hiprt_pt_render(hiprt_pt pt, uint32_t frame, void *stream);

/* Frame-level autotuner of the batched schedule (no reference counterpart;
 * the reference's CUDA tracers are compiled with fixed settings,
 * rt/cuda/trace-kernels.cu:5-10, rt/cuda/base.h:10-18).  Renders trial frames
 * of the caller's camera and path parameters (`base`; its tracer, stack,
 * samples and park fields are ignored) with candidate configurations and
 * returns the one that traces the most rays per second:
 *   1. tracer (speculative, dynamic) x layout (bvh2_scene; bvh4_scene when not
 *      NULL) x stack configuration (0 and 3; 4 on the binary tree), one
 *      sample per pixel;
 *   2. samples in flight 2, 4, ... up to max_samples for the winner of 1 and
 *      for the binary tree with stack configuration 4;
 *   3. parked leaves 1..3 for the winner.
 * Each candidate renders one warm-up and `trial_frames` (0: 2) timed frames
 * on its own path tracer.  Images do not depend on the choice.  Synchronous. */
typedef struct {
    int32_t tracer;
    int32_t stack_config;
    int32_t samples;
    int32_t bvh_format;     /* HIPRT_BVH2: bvh2_scene, HIPRT_BVH4: bvh4_scene */
    int32_t park_leaves;
    float mrays_per_s;      /* on the trial frames */
    int32_t candidates;
} hiprt_pt_tune_result;

int hiprt_pt_autotune(hiprt_scene bvh2_scene, hiprt_scene bvh4_scene, const float *cam,
                      const hiprt_pt_params *base, int32_t max_samples, int32_t trial_frames,
                      hiprt_pt_tune_result *out);
/* Live rays traced per level in the last frame (synchronizes the device). */
int hiprt_pt_ray_counts(hiprt_pt pt, uint32_t *counts, int32_t n_levels);
/* Live rays traced per level summed over every frame since creation
 * (synchronizes the device). */
int hiprt_pt_ray_totals(hiprt_pt pt, uint64_t *totals, int32_t n_levels);
/* Runs frame `frame` twice outside the graph (not with regen_pool): once timing each level's trace
 * launch with HIP events on `stream` (ms_per_level), once with the counting
 * trace variant (fetch_stats: inner-node fetches and triangle tests, 2 per
 * level).  rays_per_level receives the live rays of each level.  Both frames
 * add to the radiance accumulator.  Synchronizes the device. */
int hiprt_pt_profile(hiprt_pt pt, uint32_t frame, void *stream, float *ms_per_level, uint64_t *fetch_stats,
                     uint32_t *rays_per_level, int32_t n_levels);
/* Device memory the path tracer holds (its path slots, queues, sort
 * scratch, accumulators), in bytes; the scene is not counted. */
int hiprt_pt_memory(hiprt_pt pt, uint64_t *bytes);
/* Accumulated radiance, w*h*3 floats (synchronizes the device). */
int hiprt_pt_image(hiprt_pt pt, float *rgb);
int hiprt_pt_reset(hiprt_pt pt, void *stream);
int hiprt_pt_destroy(hiprt_pt pt);

/* -------------------------------------------------------- closest-hit adjoint
 * Backward pass of the closest hit with the hit set frozen: for every ray
 * with a valid hit in `hits` (the forward result; only t and id are read),
 * the intersection (t, beta, gamma) is recomputed from its triangle's
 * vertices and the upstream gradients dL/dt, dL/dbeta, dL/dgamma are pulled
 * back to the ray and the triangle's vertices.  No reference counterpart.
 *   d_grad_vertices  (n_vert x 3) floats, ACCUMULATED (+=), may be NULL
 *   d_grad_origin, d_grad_dir  (n x 3) floats, overwritten (zero for misses), may be NULL
 *   d_grad_corners   (n x 9) floats: dL/dA, dL/dB, dL/dC per ray, may be NULL
 * Needs a scene committed from a host scene. */
int hiprt_adjoint_closest(hiprt_scene sc, const hiprt_rays *rays, const hiprt_hits *hits, uint32_t n,
                          const float *d_grad_t, const float *d_grad_u, const float *d_grad_v, float *d_grad_vertices,
                          float *d_grad_origin, float *d_grad_dir, float *d_grad_corners, void *stream);
/* Vertex and triangle counts of a committed scene (for sizing gradients). */
int hiprt_scene_counts(hiprt_scene sc, uint32_t *n_vert, uint32_t *n_tris);

#ifdef __cplusplus
}
#endif

#endif /* ILLUM_HIPRT_H */
