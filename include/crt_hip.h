/*
 * crt_hip.h -- C ABI of the MI355X ray-tracing core (libcrt_hip.so).
 *
 * Plain C: POD structs, raw pointers and sizes, int status codes.  Every
 * entry point returns CRT_OK (0) or an error code; crt_last_error() returns a
 * thread-local message for the last failure.  No entry point aborts the
 * process on bad input.
 *
 * Precision contract: poses, camera rays and outputs are FP64 at this
 * boundary.  Inside, each ray is generated in FP64, its origin is moved in
 * FP64 to the entry point of the scene's bounding sphere ("rebasing"), and
 * only then cast to FP32 for traversal and the watertight triangle test.
 * Distances are returned as FP64 (rebase distance + FP32 local distance).
 *
 * Memory: unless a flag says otherwise, output pointers are HOST pointers and
 * the call copies results back before returning.  With CRT_OUT_DEVICE the
 * output pointers are device pointers (hipMalloc of the HIP runtime this
 * library is linked against) and, when `timing` is NULL, the call only
 * enqueues work on `stream` (hipStream_t, may be NULL for the null stream).
 *
 * Threads and streams: each scene keeps one grow-only device workspace per
 * stream.  Calls on one stream of a scene run one at a time (that stream's
 * lock) and are ordered by the stream; calls on different streams use
 * different workspaces and run concurrently on the host and overlap on the
 * device (e.g. consecutive frames on two streams, or two render batches
 * from two threads).  Different scenes, and the request server's workers
 * (own workspaces and streams), run concurrently too.
 *
 * Reference interfaces replaced (ceres-raytracer v0.6a4 paths):
 *   crt_scene_create/build  <- BodyFixedGroup ctor: transform + SweepSAH + PRO + NLO
 *                              (src/crt/rendering_body_fixed/body_fixed_group.hpp:45-89)
 *   crt_pass                <- get_inetersections / get_instances / get_normals
 *                              (src/crt/passes.hpp:13-198) via BodyFixedGroup::
 *                              intersection_pass/instance_pass/normal_pass
 *                              (body_fixed_group.hpp:111-124, src/pybind.cpp:527-578)
 *   crt_trace_rays          <- SingleRayTraverser::traverse + Closest/AnyPrimitiveIntersector
 *                              (lib/bvh/single_ray_traverser.hpp:145-163,
 *                               lib/bvh/primitive_intersectors.hpp:33-76)
 *   crt_lidar_batch         <- do_lidar / do_batch_lidar + SimpleLidar::cast_rays /
 *                              batch_cast_rays (src/crt/do_lidar.hpp:13-155,
 *                              src/crt/lidars/simple_lidar.hpp:42-96)
 *   crt_render              <- do_render + unidirectional (src/crt/do_render.hpp:27-141,
 *                              src/crt/path_tracing/unidirectional.hpp:11-96)
 *   crt_flash_lidar, crt_radiation_pressure, crt_sun_pass: no reference
 *                              counterpart (new, additive API).
 */
#ifndef CRT_HIP_H
#define CRT_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define CRT_OK               0
#define CRT_ERR_INVALID      1   /* bad argument (maps to ValueError)        */
#define CRT_ERR_TYPE         2   /* unknown camera/light/lidar type          */
#define CRT_ERR_OOM          3   /* device allocation failed (MemoryError)   */
#define CRT_ERR_HIP          4   /* HIP runtime error (RuntimeError)         */
#define CRT_ERR_STATE        5   /* scene not built / wrong call order       */
#define CRT_ERR_ADMISSION    6   /* serving: request exceeds memory budget   */
#define CRT_ERR_CANCELLED    7   /* serving: request cancelled before it ran  */
#define CRT_ERR_TIMEOUT      8   /* serving: wait timed out (ticket still valid) */

/* ---- flags ------------------------------------------------------------- */
#define CRT_OUT_DEVICE       1u  /* output (and input ray) pointers are device pointers */
#define CRT_ANY_HIT          2u  /* crt_trace_rays: stop at the first hit (shadow query) */

/* ---- light types ------------------------------------------------------- */
#define CRT_LIGHT_POINT        0 /* src/crt/lights/point_light.hpp        */
#define CRT_LIGHT_AREA         1 /* src/crt/lights/area_light.hpp         */
#define CRT_LIGHT_DIRECTIONAL  2 /* new: sun at infinity                  */

typedef struct crt_scene crt_scene;   /* opaque: device mesh + BVH, resident in HBM */

/* Pinhole camera (src/crt/cameras/simple_camera.hpp:14-27,61-81).
 * rotation is row-major R; camera rays are rotated by R^T like the reference. */
typedef struct {
    double position[3];
    double rotation[9];
    double focal_length;
    double resolution[2];     /* [x, y] = [W, H] */
    double sensor_size[2];
    int32_t z_positive;
    int32_t pad_;
} crt_camera;

typedef struct {
    int32_t type;             /* CRT_LIGHT_* */
    int32_t pad_;
    double intensity;
    double position[3];
    double rotation[9];       /* row-major R (area light samples use R, area_light.hpp:63) */
    double size[2];           /* area light side lengths */
    double direction[3];      /* directional: unit vector from the scene towards the light */
} crt_light;

#define CRT_BUILD_PLOC 1u     /* binary tree by locally-ordered clustering (PLOC) whatever the size:
                                 lower SAH, ~3x the build time */
#define CRT_BUILD_GREEDY_WIDE 2u /* collapse the binary tree into 8-wide nodes greedily (largest box
                                    first) instead of by SAH-optimal dynamic programming (default) */
#define CRT_BUILD_LBVH 4u     /* Morton-order LBVH binary tree whatever the size */
#define CRT_BUILD_PLOC_AUTO_MAX 131072u   /* without PLOC / LBVH: PLOC up to this many triangles */

typedef struct {
    uint32_t max_leaf_size;   /* triangles per leaf, 1..8 (default 8)  */
    uint32_t flags;           /* CRT_BUILD_*; 0 = PLOC binary tree for scenes of at most
                                 CRT_BUILD_PLOC_AUTO_MAX triangles, LBVH (Morton radix tree) for
                                 larger ones, then the SAH-optimal collapse */
} crt_build_opts;

typedef struct {
    uint32_t n_tris;
    uint32_t n_wide_nodes;
    uint32_t n_leaves;
    uint32_t max_depth;       /* wide levels, root = 1 */
    uint32_t stack_capacity;  /* per-ray traversal stack entries */
    uint32_t pad_;
    double   build_ms;        /* device time of the whole build */
    double   sah_cost;        /* SAH cost of the wide BVH (node cost 1, triangle cost 1) */
    double   bounds[6];       /* xmin,ymin,zmin,xmax,ymax,zmax */
    double   sphere[4];       /* bounding sphere centre xyz + radius used for rebasing */
    uint64_t device_bytes;    /* HBM held by the scene (mesh + BVH) */
    double   self_offset;     /* secondary-ray origin offset along the geometric normal:
                                 max(1e-4, 1e-6 x the scene's half extent), the reference's
                                 1e-4 (unidirectional.hpp:61) at unit scale */
    double   refit_ms;        /* device time of the last crt_scene_refit (0: built) */
} crt_bvh_info;

/* Kernel timings of the last call, measured with hipEvents on the launch stream. */
typedef struct {
    double total_ms;          /* first launch start .. last launch end          */
    double trace_ms;          /* dominant traversal kernel (closest hit)        */
    double shadow_ms;         /* any-hit kernel, if any                         */
    uint64_t rays_traced;     /* closest-hit rays                               */
    uint64_t shadow_rays;     /* any-hit rays                                   */
} crt_timing;

typedef struct {
    int32_t  min_samples;
    int32_t  max_samples;
    double   noise_threshold;
    int32_t  num_bounces;
    uint32_t pad_;
    uint64_t seed;            /* counter-based RNG key: results depend only on
                                 (seed, pixel, sample, bounce), not on batching */
} crt_render_opts;

/* ---- device / errors --------------------------------------------------- */
int         crt_device_init(int device);
const char* crt_last_error(void);
int         crt_version(void);

/* Kernel timing log (benchmarks): while enabled, crt_sun_pass records a
 * hipEvent pair around each of its kernels (k_primary_sun, k_shadow_sun) on
 * the launch stream without synchronising, so kernels can be timed inside a
 * timed region where streams overlap.  crt_timing_log_read waits for the
 * logged events of one kernel, returns their count and mean duration and
 * drops them from the log. */
int crt_timing_log_enable(int on);
int crt_timing_log_read(const char* kernel, uint64_t* count, double* mean_ms);

/* ---- scene ------------------------------------------------------------- */
/* Upload a merged triangle mesh (all entities already transformed to the
 * scene frame) and build the BVH on the GPU.
 *   vertices        : n_vertices x 3 float
 *   indices         : n_tris x 3 uint32 (vertex indices)
 *   tri_instance    : n_tris uint32 instance ids (1-based, like src/pybind.cpp:110-120)
 *   vertex_normals  : n_vertices x 3 float or NULL (then smooth shading falls back to flat)
 *   instance_color  : n_instances x 3 float (RGB albedo), may be NULL (grey 1,1,1)
 *   instance_smooth : n_instances uint8, may be NULL (flat)                              */
int crt_scene_create(const float* vertices, uint32_t n_vertices,
                     const uint32_t* indices, uint32_t n_tris,
                     const uint32_t* tri_instance,
                     const float* vertex_normals,
                     uint32_t n_instances,
                     const float* instance_color,
                     const uint8_t* instance_smooth,
                     const crt_build_opts* opts,
                     crt_scene** out_scene);
/* The same with FP64 vertices (n_vertices x 3 double), as the reference keeps
 * its transformed geometry (Scalar = double, src/pybind.cpp:30;
 * src/crt/transform.hpp:42-85).  The scene-frame origin is chosen in FP64
 * and subtracted before the single rounding to FP32, so an entity posed far
 * from the world origin keeps the FP32 resolution of its own extent. */
int crt_scene_create_f64(const double* vertices, uint32_t n_vertices,
                         const uint32_t* indices, uint32_t n_tris,
                         const uint32_t* tri_instance,
                         const float* vertex_normals,
                         uint32_t n_instances,
                         const float* instance_color,
                         const uint8_t* instance_smooth,
                         const crt_build_opts* opts,
                         crt_scene** out_scene);
/* ---- materials (src/crt/materials/material.hpp:90-184) -------------------
 * The reference's Entity hard-codes ColoredLambertianMaterial
 * (src/crt/rendering_dynamic/entity.hpp:51-53) and never reaches its other
 * materials; this additive call gives each instance one of them.  Without it
 * every instance is ColoredLambertian with its instance colour.          */
#define CRT_MAT_LAMBERTIAN      0u  /* ColoredLambertianMaterial: colour (-L.n)          */
#define CRT_MAT_TEX_LAMBERTIAN  1u  /* TexturedLambertianMaterial: map(uv) (-L.n)       */
#define CRT_MAT_BLINN_PHONG     2u  /* TexturedBlinnPhongMaterial: clamp(map (ka + kd diffuse)
                                       + ks pow(n.h, alpha)), cosine-sampled bounce       */
#define CRT_MAT_MIRROR          3u  /* MirrorMaterial: no direct light, bounce d - 2 n (d.n) */

typedef struct {
    uint32_t type;            /* CRT_MAT_* */
    int32_t  texture;         /* colour map: index into `textures`, or -1 for the constant `color` */
    int32_t  spec_texture;    /* Blinn-Phong ka / kd / ks map (texel R, G, B): index, or -1 for `coeffs` */
    float    alpha;           /* Blinn-Phong exponent (the reference's default is 24) */
    float    color[3];        /* constant colour map */
    float    coeffs[3];       /* constant ka, kd, ks */
} crt_material;

typedef struct {
    uint32_t width, height;
    const uint8_t* rgba;      /* width x height RGBA8, row-major, row 0 first (lodepng's layout) */
} crt_texture;

/* materials: one per instance (n_materials == the scene's instance count,
 * instance id i uses materials[i - 1]); vertex_uv: n_vertices x 2 floats or
 * NULL (all zero, as the reference's OBJ loader leaves them); textures are
 * copied to the device.  Texel lookup is the reference's nearest texel at
 * (size_t)(u W + 0.5), (size_t)(v H + 0.5), clamped to the image.  Not to be
 * called while other calls on the scene run; it first waits for the work
 * already enqueued on the device. */
int crt_scene_set_materials(crt_scene* scene, uint32_t n_materials, const crt_material* materials,
                            const float* vertex_uv, uint32_t n_textures, const crt_texture* textures);
/* ---- refit for moving geometry (dynamic scenes) --------------------------
 * The reference rebuilds its SweepSAH + PRO BVH on every call of the
 * dynamic API (src/crt/rendering_dynamic/render.hpp:19-62,
 * src/crt/passes.hpp:200-357, simulate_lidar.hpp:7-59) and ships a
 * refitter (lib/bvh/hierarchy_refitter.hpp:11-43).  Here: new vertex
 * positions (FP64, n_vertices x 3, same order and triangles as at creation)
 * and optionally new vertex normals; the tree keeps its topology and every
 * box is recomputed bottom-up on the GPU (unmoved vertices give the built
 * nodes bit for bit).  When the refitted SAH exceeds max_sah_growth (<= 0:
 * 1.5) times the SAH of the last full build the BVH is rebuilt instead
 * (*rebuilt = 1, may be NULL).  Closest hits equal those of a scene created
 * from the new vertices except which triangle wins an exact tie.  Not to be
 * called while other calls on the scene run; it first waits for the work
 * already enqueued on the device (e.g. device-output calls not yet done). */
int crt_scene_refit(crt_scene* scene, const double* vertices, const float* vertex_normals, double max_sah_growth,
                    uint32_t* rebuilt);
int  crt_scene_info(const crt_scene* scene, crt_bvh_info* info);
void crt_scene_destroy(crt_scene* scene);

/* ---- scene blobs (new; no reference counterpart: the reference keeps its
 * BVH cache in memory only, src/crt/rendering_body_fixed/body_fixed_group.hpp:39)
 * crt_scene_serialize: with buffer == NULL stores the blob size in *size;
 * otherwise writes the blob (mesh in the scene frame, shading data, both BVH
 * layouts, checksum) into buffer, which must hold *size bytes.
 * crt_scene_deserialize: validates the blob (magic, version, sizes, checksum,
 * node references, tree order and depth, index ranges) before any device
 * memory is used, then makes the scene resident without a rebuild.        */
int crt_scene_serialize(const crt_scene* scene, void* buffer, uint64_t* size);
int crt_scene_deserialize(const void* buffer, uint64_t size, crt_scene** out_scene);

/* ---- camera passes (src/crt/passes.hpp:13-198) --------------------------
 * One closest-hit ray per pixel at integer pixel coordinates.  Outputs are
 * row-major (H, W): xyz / normal f64 [H][W][3], instance u32 [H][W]; misses
 * give zeros, like the reference.  Any output may be NULL.  `depth` (f64
 * [H][W], distance camera->hit, 0 on miss) and `tri_id` (int32, -1 on miss)
 * are additive outputs.                                                      */
int crt_pass(crt_scene* scene, const crt_camera* camera,
             double* xyz, double* normal, uint32_t* instance,
             double* depth, int32_t* tri_id,
             uint32_t flags, void* stream, crt_timing* timing);

/* ---- general ray queries ----------------------------------------------- *
 * n rays given in FP64: origins/dirs n x 3, tmin/tmax n (NULL -> 0 / +inf).
 * Closest hit: t (in units of |dir|), tri (original triangle index, -1 miss),
 * u, v (reference barycentrics: hit = u p1 + v p2 + (1-u-v) p0).  With
 * CRT_ANY_HIT only `t` (any hit distance, -1 if none) and `tri` (>=0 if hit)
 * are meaningful.  Any output may be NULL.                                    */
int crt_trace_rays(crt_scene* scene, uint64_t n,
                   const double* origins, const double* dirs,
                   const double* tmin, const double* tmax,
                   double* t, int32_t* tri, double* u, double* v,
                   uint32_t flags, void* stream, crt_timing* timing);

/* Traversal statistics <- SingleRayTraverser::traverse(ray, intersector,
 * Statistics&) (lib/bvh/single_ray_traverser.hpp:130-160): the rays of
 * crt_trace_rays (tmin 0, tmax +inf; CRT_ANY_HIT for any hit) traced by the
 * same lane traversal, with per ray in `counts` (n x 3 uint32, host) the
 * wide nodes visited, the child boxes tested (valid children of the visited
 * nodes; the reference's traversal_steps test two boxes each) and the
 * triangles tested (a leaf counts all its triangles, like the reference's
 * intersections).  t and tri as crt_trace_rays (may be NULL).  Host arrays.  */
int crt_trace_stats(crt_scene* scene, uint64_t n, const double* origins, const double* dirs,
                    uint32_t flags, uint32_t* counts, double* t, int32_t* tri, void* stream);

/* ---- primary + directional-sun shadow pass (new) ------------------------ *
 * The camera's closest hits plus one any-hit shadow ray per hit towards the
 * sun direction (unit vector from the scene towards the sun).  Shadow rays
 * are compacted on the device.  Outputs (any may be NULL): depth f64 [H][W]
 * (0 on miss), tri_id int32 [H][W], lit uint8 [H][W] (1 = hit and sunlit).  */
int crt_sun_pass(crt_scene* scene, const crt_camera* camera, const double sun_dir[3],
                 double* depth, int32_t* tri_id, uint8_t* lit,
                 uint32_t flags, void* stream, crt_timing* timing);

/* ---- path-traced render (src/crt/do_render.hpp:27-141) -----------------
 * Lambertian multi-bounce path tracing with next-event estimation, adaptive
 * sampling (min/max samples, noise threshold) and the reference's weights.
 * Outputs (any may be NULL): rgba f32 [H][W][4] radiance (alpha 1), rgba_u8
 * [H][W][4] = trunc(clamp(256 x, 0, 255)) like the reference's uint8 image,
 * spp_used int32 [H][W].  Any number of lights (the paths of a pass shrink
 * with many lights so their shadow records stay within 4 GiB).            */
int crt_render(crt_scene* scene, const crt_camera* camera, const crt_light* lights, uint32_t n_lights,
               const crt_render_opts* opts, float* rgba, uint8_t* rgba_u8, int32_t* spp_used,
               uint32_t flags, void* stream, crt_timing* timing);

/* ---- batched render (new) ------------------------------------------------
 * n_cams cameras of one scene in one set of launches.  Camera c uses
 * max_samples[c] (NULL: opts->max_samples) and RNG seed seeds[c] (NULL:
 * opts->seed + c); its image is bit-for-bit the crt_render image of that
 * camera with that seed.  Outputs are packed in camera order, each camera
 * row-major [H][W][4] (rgba, rgba_u8) or [H][W] (spp_used).
 * Knobs (crt_batch_opts, zeros = defaults): samples per pass, hipGraph replay
 * of a pass (non-adaptive renders only), persistent traversal kernels with
 * blocks_per_cu x CUs one-wave blocks claiming `grab` rays per atomic.     */
typedef struct {
    uint32_t samples_per_pass;
    uint32_t use_graph;
    uint32_t persistent;
    uint32_t blocks_per_cu;
    uint32_t grab;
    uint32_t extend_mode;     /* closest-hit extension rays: 1 = an octet of lanes per ray,
                                 2 = one lane per ray, 3 = one lane per ray for camera rays and an
                                 octet per ray after the first bounce; 0 (default) = 3 for passes
                                 of at least 2^20 paths, 1 for smaller ones */
    uint32_t separate_kernels; /* 0: fused bounce kernels (trace + shade, shadow + accumulate;
                                 camera rays generated in the lane trace kernel); 1: one kernel
                                 per stage; 2: fused, except octet-traced bounces, which use
                                 separate trace and shade kernels.  Persistent traversal always
                                 uses separate kernels.  All give bit-identical images. */
    uint32_t coherence_sort;  /* 1: order each bounce's extension rays (after the first) by the cell
                                 of their origin before the fused trace kernels; 0 (default): queue
                                 order, which already follows the pixel tiles of the paths.  The
                                 image is the same either way. */
    uint32_t rays_per_octet;  /* octet bounce kernels: paths one octet traces in turn (1..8); 0 = by
                                 pass size (8 from 2^21 paths, 4 from 2^20, 2 from 2^19, else 1) */
    uint32_t rays_per_quad;   /* shadow kernel: live paths one quad tests in turn (1..4); 0 = by pass
                                 size (4 from 2^21 paths, 2 from 2^20, else 1) */
    uint32_t tile_width;      /* camera rays of a wave: a tile of 64 pixels this wide (4, 8, 16 or
                                 32; 0 = 8, i.e. 8x8) */
    uint32_t fixed_share;     /* fused bounce kernels: 0 (default) = a wave's octets (quads) take
                                 its paths one at a time from an LDS counter wherever the extra
                                 LDS still lets a CU hold 32 waves (BVH stacks of at most 64
                                 entries), else each takes a fixed share of R paths; 1 = always a
                                 fixed share.  The image is the same either way. */
} crt_batch_opts;
int crt_render_batch(crt_scene* scene, uint32_t n_cams, const crt_camera* cams, const int32_t* max_samples,
                     const uint64_t* seeds, const crt_light* lights, uint32_t n_lights,
                     const crt_render_opts* opts, const crt_batch_opts* bopts, float* rgba, uint8_t* rgba_u8,
                     int32_t* spp_used, uint32_t flags, void* stream, crt_timing* timing);

/* ---- LiDAR (src/crt/do_lidar.hpp:13-155) ------------------------------
 * One ray per pose along +z (z_positive) or -z rotated by R^T; range =
 * |hit - position| (0 on a miss).  positions n x 3, rotations n x 9
 * (row-major R per pose), ranges n.  A single-pose call is the reference's
 * simulate_lidar; n poses is batch_simulate_lidar without its 100000 cap.   */
int crt_lidar_batch(crt_scene* scene, uint64_t n, const double* positions, const double* rotations,
                    int32_t z_positive, double* ranges, uint32_t flags, void* stream, crt_timing* timing);

/* ---- flash LiDAR (new) --------------------------------------------------
 * Beam grid = pinhole camera `beams`; per beam the boresight ray ("single"
 * return) plus `sub_rays` rays with Gaussian angular offsets of std `sigma`
 * (rad) per tangent axis (counter RNG keyed by seed, beam, sub-ray).
 * Outputs [H][W] (any may be NULL): single, first (min), last (max) range over
 * the sub-rays that hit (0 if none), n_returns = number of sub-ray hits.     */
typedef struct {
    crt_camera beams;
    uint32_t sub_rays;        /* <= 63 (a beam's rays share one wavefront) */
    uint32_t pad_;
    double sigma;
    uint64_t seed;
} crt_flash_opts;
int crt_flash_lidar(crt_scene* scene, const crt_flash_opts* opts, double* single, double* first, double* last,
                    uint8_t* n_returns, uint32_t flags, void* stream, crt_timing* timing);

/* ---- radiation pressure / flux (new) --------------------------------------
 * Parallel rays on a grid x grid lattice over the square circumscribing the
 * disk of radius `radius` around `center` (radius <= 0: the scene's bounding
 * sphere), perpendicular to sun_dir (unit vector towards the sun).  Per hit
 * on a sun-facing triangle with flat outward normal n, cos = n.s:
 *   dF = pressure * A_ray * [(1 - rho_s) (-s) - 2 (rho_s cos + rho_d / 3) n]
 *   dT = (x_hit - ref_point) x dF
 * Per-ray terms FP32, compensated FP64 reduction.  force/torque in N, N m
 * for pressure in N/m^2 and lengths in m.                                   */
typedef struct {
    double sun_dir[3];
    double pressure;
    double rho_s;
    double rho_d;
    uint32_t grid;
    uint32_t pad_;
    double center[3];
    double radius;
    double ref_point[3];
} crt_flux_opts;
int crt_radiation_pressure(crt_scene* scene, const crt_flux_opts* opts, double force[3], double torque[3],
                           uint64_t* n_hits, void* stream, crt_timing* timing);

/* Allocation-free variant of crt_radiation_pressure.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc/hipFree or allocating routine: its scratch (block
 * partials and compensated slices, 3.7 MB) is a static device array of the
 * library, one per GPU, held until the process ends.  grid must be in
 * 1..2048 (larger grids: CRT_ERR_INVALID).  out7 is a DEVICE pointer to 7
 * doubles: force xyz, torque xyz, number of hit rays; the call only enqueues
 * work on `stream` (no synchronisation) and needs the scene's device current.
 * Results are bit-identical to crt_radiation_pressure.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (different
 * streams or threads before the previous call's kernels finished) race. */
int crt_radiation_pressure_noalloc(crt_scene* scene, const crt_flux_opts* opts, double* out7, void* stream);

/* ---- serving (new) ---------------------------------------------------------
 * A server keeps scenes resident, admits requests by device-memory budget and
 * runs them on n_workers streams; each worker batches every queued request
 * that can share its launches (continuous batching: the renders of one scene
 * and bounce count in one set of launches, a scene's flash-LiDAR scans in one
 * launch, its flux grids in one launch sequence).  batching = 0 is the plain
 * baseline: one worker, one request per launch sequence.  Outputs are host
 * buffers owned by the caller until crt_server_wait returns.  Every option is
 * validated at submission (bad requests are refused there, never inside a
 * batch); a batch that fails at run time is rerun request by request, so a
 * failure reaches only its own request.  The memory budget covers the
 * admitted requests and the device workspaces the workers keep between
 * batches.  Results depend only on the request (its seed), never on the mix. */
#define CRT_REQ_RENDER 0u
#define CRT_REQ_LIDAR  1u
#define CRT_REQ_FLUX   2u

typedef struct crt_server crt_server;

typedef struct {
    uint64_t memory_budget;     /* bytes of working memory admitted requests may hold (0: 16 GiB) */
    uint32_t n_workers;         /* concurrent batches / streams (0: 3)          */
    uint32_t max_batch;         /* requests per batch (0: 64)                   */
    uint64_t max_batch_paths;   /* camera paths (pixels x spp) per render batch; LiDAR rays per scan batch */
    uint32_t batching;          /* 0: plain per-request launches                */
    uint32_t pad_;
    crt_batch_opts render;      /* launch parameters of every render launch (zeros = defaults) */
} crt_server_opts;

typedef struct {
    uint32_t kind;              /* CRT_REQ_*                                        */
    uint32_t scene;             /* id from crt_server_add_scene                     */
    crt_camera camera;          /* render camera or LiDAR beam grid                 */
    int32_t spp;                /* render samples per pixel                         */
    int32_t bounces;            /* render bounces                                   */
    uint64_t seed;              /* render / LiDAR RNG seed                          */
    uint32_t n_lights;          /* render lights                                    */
    uint32_t sub_rays;          /* LiDAR sub-rays per beam                          */
    crt_light lights[2];        /* the lights when light_list is NULL (n_lights <= 2) */
    double sigma;               /* LiDAR sub-ray angular sigma (rad)                */
    crt_flux_opts flux;         /* radiation-pressure request                       */
    float* rgba;                /* render output [H][W][4] f32                      */
    double* lidar;              /* LiDAR output: single, first, last, each [H][W]    */
    double* force_torque;       /* flux output: force xyz, torque xyz               */
    const crt_light* light_list;/* n_lights lights, any number (copied at submission), or NULL */
} crt_request;

typedef struct {
    uint64_t completed;
    uint64_t rejected;
    uint64_t rays;              /* closest-hit + shadow rays traced                 */
    uint64_t batches;
    double p50_ms, p99_ms, mean_ms;   /* submit -> completion latency           */
    double wall_s;              /* first submission .. last completion              */
    double mrays_per_s;         /* rays / wall_s                                    */
    uint64_t cancelled;         /* requests cancelled while queued                  */
} crt_server_stats;

int  crt_server_create(const crt_server_opts* opts, crt_server** out);
int  crt_server_add_scene(crt_server* server, crt_scene* scene, uint32_t* scene_id);
int  crt_server_submit(crt_server* server, const crt_request* request, uint64_t* ticket);
int  crt_server_wait(crt_server* server, uint64_t ticket, double* latency_ms);
/* crt_server_wait with a time limit: CRT_ERR_TIMEOUT if the request has not
 * finished within timeout_ms (the ticket stays valid; wait again later). */
int  crt_server_wait_for(crt_server* server, uint64_t ticket, double timeout_ms, double* latency_ms);
/* Cancel a request that is still queued: it never runs, and its wait returns
 * CRT_ERR_CANCELLED.  A request already running or finished is not cancelled
 * (CRT_ERR_STATE); its wait returns its result as usual. */
int  crt_server_cancel(crt_server* server, uint64_t ticket);
int  crt_server_drain(crt_server* server);
int  crt_server_stats_get(crt_server* server, crt_server_stats* stats);
/* Start the statistics afresh (counts, latencies, the wall-clock window),
 * for example after a warm-up stream; the server's workspaces and scenes stay.
 * Only while nothing is queued or running (CRT_ERR_STATE otherwise). */
int  crt_server_stats_reset(crt_server* server);
void crt_server_destroy(crt_server* server);

#ifdef __cplusplus
}
#endif
#endif /* CRT_HIP_H */
