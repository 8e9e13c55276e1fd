/*
 * walkray.h -- C ABI of libwalkray, the MI355X (gfx950) implementation of the
 * WalkRay path-tracing hot path.
 *
 * Conventions (kept from the reference, scene.cpp:102-113, kdtree.cpp:300-305):
 *   - a miss is prim = WR_NONE and t = -1;
 *   - object ids are the order of wr_scene_add_mesh calls (Scene::objects);
 *   - triangle ids ("prim") are global in that order: object 0's triangles
 *     first, each object's in its own order (Scene::buildKDTree, scene.cpp:20-44).
 * Every function returns 0 on success or a negative WR_E* code; the message is
 * in wr_last_error(ctx) (or wr_last_error(NULL) for errors before a context
 * exists).  A context is one HIP device plus one stream and is not thread-safe;
 * use one per host thread.  Pointers named d_* are device pointers (caller
 * owned, e.g. from wr_device_alloc); all other pointers are host memory.
 */
#ifndef WALKRAY_H
#define WALKRAY_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WR_NONE 0xFFFFFFFFu
#define WR_OK 0
#define WR_EINVAL (-1)
#define WR_EHIP (-2)
#define WR_ESTATE (-3)
#define WR_ENOMEM (-4)
#define WR_EBUSY (-5)

typedef struct wr_ctx wr_ctx;
typedef struct wr_scene wr_scene;

/* KDTree::Ray (kdtree.h:12-15): 24 B. */
typedef struct {
    float o[3];
    float d[3];
} wr_ray;

/* The LFJ::Ray fields the NoSelf condition reads (ray.h:13-26, noself.cpp:7-11):
 * 32 B.  inside_obj / contact_obj are object ids or WR_NONE (NULL). */
typedef struct {
    float o[3];
    float d[3];
    uint32_t inside_obj;
    uint32_t contact_obj;
} wr_path_ray;

/* Closest hit: 16 B.  prim = WR_NONE and t = -1 on a miss; b1, b2 are the
 * barycentrics of Mesh::getWorldNormal (mesh.h:67-72) at o + d t.
 * prim = 0xFFFFFFFE flags a ray whose traversal stack overflowed (never
 * silently wrong; trees deeper than 64 levels are rejected at build). */
typedef struct {
    uint32_t prim;
    float t;
    float b1;
    float b2;
} wr_hit;

/* Materials of the four hot object classes (diffuseObj.h, glossyObj.h,
 * refractiveObj.h, emissiveObj.h). */
enum { WR_DIFFUSE = 0, WR_GLOSSY = 1, WR_REFRACTIVE = 2, WR_EMISSIVE = 3 };
typedef struct {
    int32_t type;
    float color[3]; /* albedo (diffuse, glossy) or emitted radiance (emissive) */
    float power;    /* Phong exponent (glossy) */
    float ior;      /* index of refraction (refractive) */
} wr_material;

/* KDTree::Strategy (kdtree.h:20). Scene always builds with BEST. */
enum { WR_LOOP = 0, WR_BEST = 1 };

typedef struct {
    uint32_t n_objects;
    uint32_t n_triangles;
    uint32_t n_nodes;       /* internal nodes of the global tree */
    uint32_t n_leaves;
    uint32_t depth;         /* internal levels of the global tree */
    uint32_t max_obj_depth; /* deepest per-object tree */
    uint64_t device_bytes;  /* device memory held by the scene */
    double build_seconds;   /* host build (all trees) */
} wr_scene_info;

/* ---- context ----------------------------------------------------------- */
/* Replaces nothing in the reference (single process, CPU). */
int wr_ctx_create(int device, wr_ctx** out);
void wr_ctx_destroy(wr_ctx* ctx);
const char* wr_last_error(const wr_ctx* ctx);
int wr_synchronize(wr_ctx* ctx);
int wr_device_alloc(wr_ctx* ctx, size_t bytes, void** d_out);
int wr_device_free(wr_ctx* ctx, void* d_ptr);
int wr_memcpy_htod(wr_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int wr_memcpy_dtoh(wr_ctx* ctx, void* dst, const void* d_src, size_t bytes);
/* Asynchronous device-to-device copy on the context stream. */
int wr_memcpy_dtod(wr_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* HIP events on the context's stream, bracketing kernel work. */
int wr_timer_start(wr_ctx* ctx);
int wr_timer_stop(wr_ctx* ctx, float* ms);

/* ---- scene ------------------------------------------------------------- */
/* Scene (scene.h:10-45): objects + the global tree + per-object trees. */
int wr_scene_create(wr_ctx* ctx, wr_scene** out);
void wr_scene_destroy(wr_scene* scene);
/* Adds one object (AbstractObject + Mesh, abstractObj.h:10-49, mesh.h:9-19) in
 * world space.  normals: per-vertex shading normals (nv * 3) or NULL (flat,
 * mesh.h:59-65).  *obj_id receives the object id. */
int wr_scene_add_mesh(wr_scene* scene, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* normals, const wr_material* mat, uint32_t* obj_id);
/* wr_scene_add_mesh with shading normals per triangle corner (nt * 9): the
 * world normals Mesh::getWorldNormal interpolates (mesh.h:59-72), for meshes
 * whose normal indices differ from their position indices. */
int wr_scene_add_mesh_corners(wr_scene* scene, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                              const float* corner_normals, const wr_material* mat, uint32_t* obj_id);
/* Scene::buildKDTree + buildObjKDTrees (scene.cpp:20-65) with
 * KDTree::build(BEST) (kdtree.cpp:242-258), flattened and uploaded.
 * WR_EINVAL for a NaN vertex coordinate in a triangle: the split rule's
 * median (std::nth_element, kdtree.cpp:179) is undefined on it. */
int wr_scene_build(wr_scene* scene, int strategy);
int wr_scene_get_info(const wr_scene* scene, wr_scene_info* info);
/* prim -> (object id, triangle id in the object): the ObjSourceInfo of
 * scene.cpp:34-36,110. */
int wr_scene_prim_source(const wr_scene* scene, uint32_t prim, uint32_t* obj, uint32_t* tri);
/* The same for n hits on the device (Scene::intersect's ObjSourceInfo out
 * parameter, scene.cpp:110, for a whole batch): d_obj[i], d_tri[i] = the
 * object and its own triangle index of d_hits[i].prim, WR_NONE for a miss or a
 * stack-overflow record.  Asynchronous on the context stream; device pointers. */
int wr_hit_sources(wr_ctx* ctx, const wr_scene* scene, const wr_hit* d_hits, uint64_t n, uint32_t* d_obj,
                   uint32_t* d_tri);
/* The scene as pass `pass` of a render with `seed` places it: what wr_render
 * does before every pass of a scene with moving objects (Scene::
 * updateSceneForMotionBlur, scene.cpp:67-80, called at pathTracer.cpp:18):
 * each moving object at its time for that pass, the reference trees rebuilt
 * exactly, the trees of objects that do not move kept, and the fast-pass
 * hierarchies of the rebuilt trees refit on the device.  *out is a built,
 * static scene of its own (queries, renders, exports; wr_scene_destroy).  No
 * counterpart in the reference beyond that call, whose scene it exposes. */
int wr_scene_motion_pass(const wr_scene* scene, uint64_t seed, uint32_t pass, wr_scene** out);
/* The fast-pass hierarchy of the global tree (obj = WR_NONE) or of object
 * obj's tree as on the device: n_wide 64-byte nodes (wr_tree_wide's layout)
 * and the root reference.  Pass NULL nodes to get the sizes only.  No
 * counterpart in the reference (its tree has no such hierarchy). */
int wr_scene_export_wide(const wr_scene* scene, uint32_t obj, uint32_t* n_wide, uint32_t* wide_root, void* nodes);
/* Canonical preorder export of the global tree (obj = WR_NONE) or of object
 * obj's tree, for comparison with another builder.  Pass NULL arrays to get
 * the sizes only. */
int wr_scene_export_tree(const wr_scene* scene, uint32_t obj, uint32_t* n_nodes, uint32_t* n_leaf_tris,
                         float* box, uint32_t* right, uint32_t* leaf_first, uint32_t* leaf_count,
                         uint32_t* tri_list);

/* Host-only (no device needed): runs the builder of wr_scene_build on one
 * triangle soup (KDTree::build(BEST), kdtree.cpp:242-258) and keeps the
 * result for wr_tree_export, in the same canonical preorder as
 * wr_scene_export_tree.  Used to check the builder on machines without a GPU. */
typedef struct wr_tree wr_tree;
int wr_tree_build(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads, wr_tree** out);
int wr_tree_export(const wr_tree* tree, uint32_t* n_nodes, uint32_t* n_leaf_tris, uint32_t* depth, float* box,
                   uint32_t* right, uint32_t* leaf_first, uint32_t* leaf_count, uint32_t* tri_list);
void wr_tree_free(wr_tree* tree);
/* Host-only: wr_tree_build for the same triangles at other vertex positions
 * (a moving object's pass), the reference tree built anew by the same rule
 * and the fast-pass hierarchy of prev refit to the new positions instead of
 * built anew: what wr_render's per-pass rebuild of moving objects does (no
 * counterpart in the reference, whose updateSceneForMotionBlur rebuilds every
 * tree, scene.cpp:67-80).  Exposed so the refit hierarchy can be checked
 * without a GPU. */
int wr_tree_build_refit(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads,
                        const wr_tree* prev, wr_tree** out);
/* Host-only: the fast-pass hierarchy the GPU traverses for this tree (no
 * counterpart in the reference; exposed so its invariants can be checked
 * without a GPU).  wide_nodes receives n_wide 64-byte nodes (wr_wnode of
 * walkray_amd/csrc/wr_layout.h); record_first[r] is the first record of the
 * reference leaf holding record r, record_tri[r] the input triangle of record
 * r.  Call with NULL arrays to get the sizes first. */
int wr_tree_wide(const wr_tree* tree, uint32_t* n_wide, uint32_t* n_records, uint32_t* wide_root, void* wide_nodes,
                 uint32_t* record_first, uint32_t* record_tri);

/* ---- ray queries ------------------------------------------------------- */
/* KDTree::intersect(ray, triID) without Condition (kdtree.cpp:349-353) for n
 * rays, as Scene::intersect uses it for camera rays (camera.cpp:52).
 * Asynchronous on the context stream.  d_counters (nullable, 3 x u64,
 * accumulated): box records fetched, triangle records tested, fallback rays. */
int wr_closest_hit(wr_ctx* ctx, const wr_scene* scene, const wr_ray* d_rays, uint64_t n, wr_hit* d_hits,
                   uint64_t* d_counters);
/* Autotuner (no counterpart in the reference).  Times wr_closest_hit on the
 * given rays (a representative sample: every call is traced in full) under
 * candidate launch configurations of the persistent kernel (refill threshold,
 * node-phase threshold, workgroups per CU, traversal stack levels in LDS, rays
 * per work-counter atomic), by coordinate descent from the
 * scene's current choice, and keeps the fastest for the scene; later
 * wr_closest_hit calls use it.  Results never depend on the configuration.
 * Synchronous; allocates 16 * n bytes of scratch while it runs. */
typedef struct {
    uint32_t refill_min;     /* idle lanes of a wave before it takes new rays (1..64) */
    uint32_t node_phase_min; /* lanes at wide nodes for a node-only step (1..65; 65 = never) */
    uint32_t blocks_per_cu;  /* resident 256-lane workgroups per CU, 0 = as many as fit */
    uint32_t lds_levels;     /* traversal stack levels kept in LDS: 4, 8 or 16 (0 = 8); deeper levels spill */
    uint32_t batch;          /* rays a wave takes from its XCD's work counter at a time, 64..4096 (0 = 256) */
    float ms;                /* one call on the rays with this configuration */
} wr_closest_hit_tuning;
int wr_autotune_closest_hit(wr_ctx* ctx, wr_scene* scene, const wr_ray* d_rays, uint64_t n,
                            wr_closest_hit_tuning* best);
int wr_scene_set_closest_hit_tuning(wr_scene* scene, const wr_closest_hit_tuning* tuning);
/* Motion blur (AnimatedTransform, motionBlur.h:6-34, motionBlur.cpp:6-87; set
 * by the loader at config.cpp:28-45): object obj's mesh, as added, is in its
 * own space and moves from transform `start` at start_time to `end` at
 * end_time (4x4, row-major: element (row, col) at [4 * row + col]).  Before
 * every pass wr_render draws the object's time uniformly in [0, 1) (counter
 * RNG: seed, pixel slot = obj, the pass, vertex 0xFFFFFFFE), places it with
 * AnimatedTransform::Interpolate, and rebuilds the trees
 * (Scene::updateSceneForMotionBlur, scene.cpp:67-80).  The built scene, and
 * every other query on it, shows the object at start_time.  Call before
 * wr_scene_build; not for served scenes. */
int wr_scene_set_motion(wr_scene* scene, uint32_t obj, float start_time, float end_time, const float* start,
                        const float* end);
/* AnimatedTransform(start, start_time, end, end_time).Interpolate(time) as a
 * host utility: out = the transform at `time`, row-major.  Host only. */
int wr_motion_interpolate(float start_time, float end_time, const float* start, const float* end, float time,
                          float* out);

/* Scene::intersect with NoSelfCondition(scene, ray) (mcRenderer.cpp:42-44,
 * noself.cpp:13-20) for n path rays. */
int wr_closest_hit_noself(wr_ctx* ctx, const wr_scene* scene, const wr_path_ray* d_rays, uint64_t n,
                          wr_hit* d_hits);
/* Scene::findInsideObject(ray, cur) (scene.cpp:115-144); cur is taken from
 * d_rays[i].contact_obj (WR_NONE = NULL). */
int wr_inside_object(wr_ctx* ctx, const wr_scene* scene, const wr_path_ray* d_rays, uint64_t n, uint32_t* d_obj);

/* ---- path tracing ------------------------------------------------------ */
/* Camera::setup arguments (camera.h:28-48, config.cpp:504-511): pinhole, no
 * depth of field. */
typedef struct {
    float pos[3];
    float focus[3];
    float up[3];
    uint32_t width;
    uint32_t height;
    float sight_dist; /* mImagePlaneDist, in pixels */
} wr_camera;

enum { WR_MEGAKERNEL = 0, WR_WAVEFRONT = 1 };
/* Render flags.  WR_RENDER_NEE: PathTracer's useNextEventEstimation branch
 * (pathTracer.cpp:39-66; the reference constructs it false, pathTracer.h:14):
 * a path that escapes or reaches the depth limit connects the origin of its
 * last traced ray, if diffuse, to a light sample (Scene::genLightSample,
 * scene.cpp:16-18) through MCRenderer::visibilityTest (mcRenderer.cpp:59-65).
 * The light sample's uniforms use vertex counter 0xFFFFFFFF. */
enum { WR_RENDER_NEE = 1 };
/* Wavefront launch modes (no counterpart in the reference; results identical
 * to the default): WR_RENDER_DIRECT launches the iteration kernels one by one
 * instead of replaying the captured graphs (for tools that cannot follow graph
 * launches, such as kernel tracers); WR_RENDER_STAGE_STATS does the same with a
 * HIP event around every kernel and the counting trace kernel, and keeps the
 * per-stage times and counts for wr_render_stage_stats.  Wavefront mode only. */
enum { WR_RENDER_DIRECT = 2, WR_RENDER_STAGE_STATS = 4 };
typedef struct {
    uint32_t spp;        /* passes to add, one sample per pixel each (PathTracer spp) */
    uint32_t max_depth;  /* MCRenderer maxPathLen (mcRenderer.h:16) */
    uint64_t seed;       /* counter-RNG key */
    uint32_t mode;       /* WR_MEGAKERNEL (one lane per pixel) or WR_WAVEFRONT */
    uint32_t wave_slots; /* wavefront path-slot pool size, 0 = automatic */
    uint32_t wave_grid;  /* wavefront workgroups per stage launch, 0 = automatic */
    uint32_t flags;      /* WR_RENDER_* */
} wr_render_params;

typedef struct {
    uint64_t paths;       /* paths traced by this call */
    uint64_t iterations;  /* wavefront iterations (0 for the megakernel) */
    float kernel_ms;      /* device time of this call (HIP events on the context stream; with moving
                             objects it spans the per-pass host rebuilds too) */
    float build_ms;       /* with moving objects: host time of the per-pass tree rebuilds and their
                             upload (Scene::updateSceneForMotionBlur); 0 for a static scene */
} wr_render_stats;

/* Film: per pixel and channel a Kahan-compensated FP32 sum of the per-pass
 * samples, its compensation term and a plain sum of squares, plus the number
 * of passes.  Replaces the running-mean vector<vec3f> of pathTracer.cpp:10,24,69;
 * mean = sum / passes.  Pixels are row-major, top row first (camera.cpp:14-15). */
typedef struct wr_film wr_film;
int wr_film_create(wr_ctx* ctx, uint32_t width, uint32_t height, wr_film** out);
void wr_film_destroy(wr_film* film);
int wr_film_clear(wr_film* film);
/* Any output pointer may be NULL.  mean_rgb = sum / passes (3 floats per pixel). */
int wr_film_read(wr_film* film, float* mean_rgb, float* sum, float* comp, float* sumsq, uint32_t* passes);
/* Restores a film read with wr_film_read (checkpoint / resume). */
int wr_film_write(wr_film* film, const float* sum, const float* comp, const float* sumsq, uint32_t passes);

/* PathTracer::renderPixels (pathTracer.cpp:7-76) for passes
 * [passes, passes + spp) of the film, accumulated into it.  stats nullable. */
int wr_render(wr_ctx* ctx, const wr_scene* scene, const wr_camera* cam, const wr_render_params* params,
              wr_film* film, wr_render_stats* stats);

/* Per-stage statistics of the context's last wr_render with
 * WR_RENDER_STAGE_STATS: device milliseconds summed over the iterations for
 * each of the eight iteration kernels (extend: trace, finish; shade: common
 * case, refraction exits; plan; generate; fold; advance), the path segments
 * traced, the refraction exits, the camera paths generated, and the fast-pass
 * records the trace stage fetched (nodes in 32 B units, 48 B triangles). */
typedef struct {
    float ms[8];
    uint64_t segments;
    uint64_t exits;
    uint64_t generated;
    uint64_t box_records;
    uint64_t tri_records;
} wr_stage_stats;
int wr_render_stage_stats(wr_ctx* ctx, wr_stage_stats* out);

/* BidirectionalPathTracer::renderPixels (bidirPathTracer.cpp:10-51) for
 * passes [passes, passes + spp) of the film: per pixel and pass an eye
 * sub-path and a light sub-path (the light-path counters of wr_light_paths,
 * keyed by the pixel), every connection with its MIS weight
 * (bidirPathTracer.cpp:53-313), light-tracing contributions splatted to the
 * pixel the camera sees them in; the pass's pixel values checked
 * (isLegalColor), vignetting-fixed and added to the film.  params: spp,
 * max_depth (maxPathLen, at most 32), seed, and wave_slots: the entries of
 * the visibility queue (connection rays traced in a batch after the
 * connections are formed; 0 = automatic, a full queue makes the remaining
 * connections trace in place); mode, wave_grid and flags must be 0.  The
 * contributions of a pixel are summed with float atomics, so the film is
 * reproducible up to their summation order.  Static scenes only. */
int wr_render_bdpt(wr_ctx* ctx, const wr_scene* scene, const wr_camera* cam, const wr_render_params* params,
                   wr_film* film, wr_render_stats* stats);

/* VCMTracer::renderPixels (vcmTracer.cpp:6-84) for passes [passes, passes +
 * spp) of the film.  Per pass s: the merging radius wr_vcm_radius(base_radius,
 * s); one light sub-path per pixel and the pixel's eye sub-path (the sub-paths
 * of wr_render_bdpt); the light points (every non-delta vertex past the first
 * of each light path, up to an emitter) after width * height default ones,
 * in path order, and the hash grid over them (width * height cells,
 * hashgrid.hxx:18-75); per pixel the vertex connections of wr_render_bdpt
 * with vcmMisWeight (vcmTracer.cpp:88-173,221-320) and the vertex merges of
 * every non-delta eye vertex before an emitter with the light points within
 * the radius (vcmTracer.cpp:175-217), merged contributions divided by the
 * pixel count; the pass's pixel values checked, vignetting-fixed and added to
 * the film.  params as for wr_render_bdpt.  base_radius: the configuration's
 * radius (config.cpp:560).  Reproducible up to the float atomics' summation
 * order.  Static scenes only. */
int wr_render_vcm(wr_ctx* ctx, const wr_scene* scene, const wr_camera* cam, const wr_render_params* params,
                  float base_radius, wr_film* film, wr_render_stats* stats);
/* The merging radius of pass s: MAX(base_radius * sqrt(powf(s + 1, alpha - 1)),
 * EPSILON) with alpha = 2/3 (vcmTracer.cpp:30, vcmTracer.h:39-40). */
float wr_vcm_radius(float base_radius, uint32_t pass);

/* PhotonMap::renderPixels (photonMap.cpp:21-148) with progressive photon
 * mapping (PPM) for passes [passes, passes + spp) of the film.  Per pass s:
 * the radius wr_vcm_radius(base_radius, s) (photonMap.cpp:47), the light
 * sub-paths, light points and hash grid of wr_render_vcm; per pixel the eye
 * "merge path" (samplePath with maxPathLen 10, photonMap.cpp:150-190) up to
 * its first emitter, whose emitted radiance it takes, or its first diffuse
 * vertex, where it takes the kernel density estimate of the photons within the
 * radius (kernel 3 (1 - d^2/r^2)^2 / pi over photon count x r^2,
 * photonMap.cpp:198-248,656-668); times the camera vertex's factor
 * (photonMap.cpp:670-674); the pass's pixel values checked, vignetting-fixed
 * and added to the film.  Every object here is a surface, and none is an
 * anisotropic glossy one, so the volume and ellipse branches do not arise.
 * params: spp, max_depth (of the light sub-paths, at most 32), seed; mode,
 * wave_slots, wave_grid and flags must be 0.  Static scenes only. */
int wr_render_ppm(wr_ctx* ctx, const wr_scene* scene, const wr_camera* cam, const wr_render_params* params,
                  float base_radius, wr_film* film, wr_render_stats* stats);

/* ---- light sub-paths and the photon hash grid ------------------------------ */
/* One Ray of a Path (ray.h:8-37) as the light-path estimators read it. */
typedef struct {
    float origin[3];
    float direction[3];
    float radiance[3];
    float origin_prob;
    float direction_prob;
    uint32_t contact_obj; /* contactObj, WR_NONE = NULL */
    uint32_t contact_tri; /* contactTriangleID (the object's own triangle index), 0 without a contact */
    uint32_t inside_obj;  /* insideObj, WR_NONE = NULL */
    uint32_t flags;       /* WR_VERTEX_DELTA: isDeltaDirection */
    uint32_t pad;
} wr_path_vertex;
enum { WR_VERTEX_DELTA = 1 };

/* The light sub-paths of BDPT, VCM and PPM: `Ray lightRay = genLightSample();
 * samplePath(lightPath, lightRay)` (bidirPathTracer.cpp:31-33,
 * vcmTracer.cpp:37-41, photonMap.cpp:61-65), for paths path0 .. path0 +
 * n_paths - 1 of one pass.  The light ray is AbstractObject::genSurfaceSample
 * (abstractObj.cpp:18-52); randomness is the counter RNG keyed by (seed, path,
 * pass, vertex) with vertices 0x80000000 (light pick and point), 0x80000001
 * (direction) and 0x80000002 + d (the scatter at depth d).  d_vertices holds
 * max_vertices records per path (max_vertices >= max_depth + 3, the longest
 * possible path); d_counts each path's vertex count.  Device pointers. */
int wr_light_paths(wr_ctx* ctx, const wr_scene* scene, uint32_t max_depth, uint64_t seed, uint32_t pass,
                   uint32_t path0, uint32_t n_paths, uint32_t max_vertices, wr_path_vertex* d_vertices,
                   uint32_t* d_counts);
/* The light points VCM / PPM keep from those paths (vcmTracer.cpp:42-55,
 * photonMap.cpp:66-86): per path in order, vertices 1.. up to the first whose
 * contact object is emissive, skipping isDeltaDirection ones; after n_leading
 * points at the origin (the reference's vector starts with one
 * value-initialised point per pixel, vcmTracer.cpp:31).  Writes at most cap
 * positions (d_pos, 3 floats each) and sources (d_src: path, vertex;
 * WR_NONE for the leading points); *n_out = the number there are.  Device
 * pointers except n_out. */
int wr_photons_collect(wr_ctx* ctx, const wr_scene* scene, const wr_path_vertex* d_vertices, const uint32_t* d_counts,
                       uint32_t n_paths, uint32_t max_vertices, uint64_t n_leading, float* d_pos, uint32_t* d_src,
                       uint64_t cap, uint64_t* n_out);
/* HashGrid (hashgrid.hxx:15-203): Reserve(n_cells) + Build(points, radius)
 * over n points (d_pos, device, read again by queries: keep it alive), and
 * Process as a range query: for each query point (device, 3 floats), the
 * points handed to the query in the reference's order (8 cells, a cell
 * reached twice through a hash collision visited twice), the first cap of
 * them in d_idx (cap per query), their number in d_counts.  A query call
 * keeps a device work list of n_queries + 1 words in the grid (allocated
 * when a call needs more than the last), so calls on one grid must not
 * run concurrently. */
typedef struct wr_hashgrid wr_hashgrid;
int wr_hashgrid_build(wr_ctx* ctx, const float* d_pos, uint64_t n, float radius, uint32_t n_cells,
                      wr_hashgrid** out);
/* Host copies of the grid: bbox (min xyz, max xyz), mCellEnds (n_cells), mIndices (n). */
int wr_hashgrid_export(const wr_hashgrid* grid, float* bbox, int32_t* cell_ends, int32_t* indices);
int wr_hashgrid_query(wr_hashgrid* grid, const float* d_queries, uint64_t n_queries, uint32_t cap, uint32_t* d_counts,
                      uint32_t* d_idx);
/* wr_hashgrid_query without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same results as wr_hashgrid_query.  The deferred-query work list is a fixed
 * array of the library instead of the grid's: WR_HASHGRID_NOALLOC_MAX_QUERIES
 * + 1 words (16 MiB) on every GPU the library runs on, held until the process
 * ends; more queries fail with WR_EINVAL.  Asynchronous on the context stream
 * (no synchronisation and no memset: the list's counter is cleared by a
 * kernel).  Every call on one GPU shares that array, so calls that overlap on
 * one GPU (from two contexts, or while another context's call is in flight)
 * race; calls on one context are ordered by its stream. */
#define WR_HASHGRID_NOALLOC_MAX_QUERIES (1u << 22)
int wr_hashgrid_query_noalloc(wr_hashgrid* grid, const float* d_queries, uint64_t n_queries, uint32_t cap,
                              uint32_t* d_counts, uint32_t* d_idx);
void wr_hashgrid_destroy(wr_hashgrid* grid);

/* ---- scene files ---------------------------------------------------------- */
/* Config::load (config.cpp:432-575): reads a reference scene configuration
 * (Config -> Scene -> GroupObject / Object / Camera, materials by path and
 * name) with its .obj / .shp shapes, unitized and transformed as the
 * reference does (mesh.cpp:5-174, mesh.h:106).  Only the four hot object
 * classes, static transforms, <motion> (wr_config_get_motion) and the pinhole
 * camera are accepted; anything else (other classes, depth of field,
 * refractive decay) fails with a message.  Host only. */
typedef struct wr_config wr_config;
typedef struct {
    uint32_t n_objects;
    int32_t spp;          /* <iter> (INT_MAX if absent); 0 when there is no <renderer> */
    int32_t max_len;      /* the loader's fixed maxPathLen, 20 (config.cpp:523); 0 without <renderer> */
    wr_camera camera;
    char renderer[32];    /* <renderer> as written ("PathTracer", "PT", ...), "" if absent */
    char scene_name[64];
    float radius;         /* <radius> (VCMTracer / PhotonMap base radius, config.cpp:560,571), 0 if absent */
} wr_config_info;
int wr_config_load(const char* path, wr_config** out);
void wr_config_free(wr_config* cfg);
int wr_config_get_info(const wr_config* cfg, wr_config_info* info);
/* Object i in Scene::objects order: world-space vertices (nv * 3), triangles
 * (nt * 3), world shading normals per triangle corner (nt * 9, only when
 * *has_normals), material.  Pass NULL arrays to get the sizes first. */
int wr_config_get_object(const wr_config* cfg, uint32_t i, uint32_t* nv, uint32_t* nt, uint32_t* has_normals,
                         float* verts, uint32_t* tris, float* corner_normals, wr_material* mat);
/* wr_scene_add_mesh_corners for every object, in order, and
 * wr_scene_set_motion for the moving ones. */
int wr_config_add_to_scene(const wr_config* cfg, wr_scene* scene);
/* Object i's <motion> (config.cpp:28-45): *moving = 0 for a static object;
 * otherwise its vertices and normals from wr_config_get_object are in its own
 * space and it moves from `start` at start_time to `end` at end_time
 * (row-major 4x4).  Any output but moving may be NULL. */
int wr_config_get_motion(const wr_config* cfg, uint32_t i, int* moving, float* start_time, float* end_time,
                         float* start, float* end);

/* Mesh::loadShape(path, normalize) (mesh.cpp:5-130) for one .obj or .shp file:
 * positions, faces, vertex normals and per-face normal indices, unitized when
 * normalize != 0 (mesh.cpp:152-165).  Host only. */
typedef struct wr_shape wr_shape;
int wr_shape_load(const char* path, int normalize, wr_shape** out);
void wr_shape_free(wr_shape* shape);
int wr_shape_get(const wr_shape* shape, uint32_t* nv, uint32_t* nt, uint32_t* nn, uint32_t* nfn, float* verts,
                 uint32_t* tris, float* normals, uint32_t* normal_tris);

/* Film::setBuffer + Film::savePFM (film.h:36-42,100-110): rgb is width x
 * height x 3 floats, row-major, top row first (wr_film_read's mean_rgb); the
 * file has the reference's header and rows bottom to top.  Host only. */
int wr_pfm_write(const char* path, const float* rgb, uint32_t width, uint32_t height);

/* Autotuner for the wavefront engine (no counterpart in the reference): times
 * spp-pass renders of the camera under candidate path-slot pool sizes (queue
 * capacity) and stage grids (workgroups per stage launch), then, at the best of
 * those, the traversal stack levels kept in LDS by the trace stage, and keeps
 * the fastest in the context; wr_render calls whose wave_slots / wave_grid are
 * 0 use it (the stack levels always).  Films never depend on the choice.
 * Synchronous. */
typedef struct {
    uint32_t wave_slots;
    uint32_t wave_grid;
    uint32_t lds_levels; /* 4, 8 or 16 */
    float ms;            /* the spp-pass render with this configuration */
} wr_render_tuning;
int wr_autotune_render(wr_ctx* ctx, const wr_scene* scene, const wr_camera* cam, uint32_t max_depth, uint32_t spp,
                       wr_render_tuning* best);
/* Sets the context's wavefront configuration directly (for example one tuned
 * earlier): wave_slots / wave_grid 0 keep the defaults, lds_levels must be 4,
 * 8 or 16 (0 = 8); ms is ignored.  No counterpart in the reference. */
int wr_set_render_tuning(wr_ctx* ctx, const wr_render_tuning* tuning);

/* ---- serving (no counterpart in the reference: RenderEngine runs one job at
 * a time from a key loop, renderEngine.cpp:72-91) ------------------------- */
/* A server keeps scenes resident, accepts render jobs at any time, admits them
 * against a device-memory budget and a job-slot limit, and renders every
 * admitted job's paths interleaved in one shared wavefront pool.  A job's film
 * is bit-identical to rendering it alone (randomness is keyed per job, pixel,
 * pass and vertex, and passes are folded in order). */
typedef struct wr_server wr_server;
typedef struct {
    uint64_t job_memory_budget; /* device bytes for job state (films + sample windows); 0 = 8 GiB */
    uint32_t path_slots;        /* shared path pool; 0 = 4,194,304 */
    uint32_t max_jobs;          /* concurrently admitted jobs, <= 64; 0 = 64 */
} wr_server_config;

typedef struct {
    uint32_t scene; /* id returned by wr_server_add_scene */
    wr_camera cam;
    uint32_t spp;
    uint32_t max_depth;
    uint32_t flags; /* WR_RENDER_* */
    uint64_t seed;
} wr_job_desc;

enum { WR_JOB_QUEUED = 0, WR_JOB_RUNNING = 1, WR_JOB_DONE = 2, WR_JOB_FAILED = 3, WR_JOB_CANCELLED = 4 };
typedef struct {
    int32_t state;
    uint32_t reserved;
    double submit_s, admit_s, done_s; /* wr_server_clock() times; 0 until reached */
    uint64_t samples;                 /* width * height * spp */
    uint64_t device_bytes;            /* job state held while admitted */
} wr_job_status;

int wr_server_create(wr_ctx* ctx, const wr_server_config* cfg, wr_server** out);
void wr_server_destroy(wr_server* srv);
/* Makes a built scene resident for jobs; the scene must outlive the server. */
int wr_server_add_scene(wr_server* srv, const wr_scene* scene, uint32_t* scene_id);
/* Queues a job (never blocks); a job whose state alone exceeds the budget fails. */
int wr_job_submit(wr_server* srv, const wr_job_desc* job, uint64_t* job_id);
/* Admits queued jobs that fit, runs one batch of wavefront iterations on every
 * admitted job, and retires finished ones.  *n_active (nullable) = admitted +
 * queued jobs left.  A job whose admission fails (device allocation) becomes
 * WR_JOB_FAILED, its buffers are freed, and admission goes on with the next
 * queued job: the step itself still succeeds. */
int wr_server_step(wr_server* srv, uint32_t* n_active);
/* Cancels a queued or running job: it becomes WR_JOB_CANCELLED at once and its
 * film and sample window are freed (return to the budget) at once; a running
 * job's paths leave the shared pool during the next wr_server_step, after
 * which its job slot is reused.  Every other job's film is unchanged (bit for
 * bit the same as without the cancelled job).  Cancelling a cancelled job does
 * nothing; a finished or failed job gives WR_ESTATE. */
int wr_job_cancel(wr_server* srv, uint64_t job_id);
/* Test hook for failure isolation: the next n admissions fail as if their
 * device allocation had failed. */
int wr_server_inject_admit_failures(wr_server* srv, uint32_t n);
int wr_job_get_status(wr_server* srv, uint64_t job_id, wr_job_status* status);
/* Film of a finished job (any pointer may be NULL), as wr_film_read. */
int wr_job_read(wr_server* srv, uint64_t job_id, float* mean_rgb, float* sum, float* comp, float* sumsq,
                uint32_t* passes);
/* Frees a finished or failed job's film. */
int wr_job_release(wr_server* srv, uint64_t job_id);
/* Seconds since the server was created (steady clock). */
double wr_server_clock(const wr_server* srv);

#ifdef __cplusplus
}
#endif
#endif
