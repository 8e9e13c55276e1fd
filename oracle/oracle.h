/*
 * oracle.h -- C ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The oracle is a plain C++ restatement of the WalkRay reference's path-tracing
 * hot path, written to follow the reference's algorithms and floating-point
 * operation order.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline: the
 * product library (walkray_amd/, include/walkray.h) never calls it.
 *
 * Every function below names the reference file:line it restates.
 */
#ifndef WALKRAY_ORACLE_H
#define WALKRAY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NONE 0xFFFFFFFFu

/* ---- KDTree (kdtree.h:9-48, kdtree.cpp) ------------------------------- */
/* Builds KDTree::build(BEST) (kdtree.cpp:242-258 -> splitBest 141-240). */
void*    orc_tree_build(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt);
void     orc_tree_free(void* tree);
/* Canonical preorder export: node count, leaf-triangle count, depth. */
void     orc_tree_stats(void* tree, uint32_t* n_nodes, uint32_t* n_leaf_tris, uint32_t* depth, uint32_t* n_leaves);
/* box[6*i] = min xyz, max xyz; right[i] = preorder index of right child or ORC_NONE for a leaf;
 * leaf_first/leaf_count index into tri_list (valid for leaves). */
void     orc_tree_export(void* tree, float* box, uint32_t* right, uint32_t* leaf_first,
                         uint32_t* leaf_count, uint32_t* tri_list);
/* KDTree::intersect (kdtree.cpp:349-353, 308-347, 279-306, 46-64), no Condition.
 * rays: 6 floats per ray (origin xyz, direction xyz).  prim = ORC_NONE, t = -1 on a miss. */
void     orc_tree_intersect(void* tree, const float* rays, uint64_t n, uint32_t* prim, float* t,
                            int nthreads);
/* Reference box test Node::BoundingBox::intersect (kdtree.cpp:46-64) on a single box. */
float    orc_box_intersect(const float* box6, const float* ray6);
/* Reference triangle test (kdtree.cpp:279-306) without Condition. */
float    orc_tri_intersect(const float* v0, const float* v1, const float* v2, const float* ray6);

/* ---- Scene (scene.cpp) with materials (diffuse/glossy/refractive/emissive) ---- */
enum { ORC_DIFFUSE = 0, ORC_GLOSSY = 1, ORC_REFRACTIVE = 2, ORC_EMISSIVE = 3 };
void*    orc_scene_create(void);
void     orc_scene_free(void* s);
/* params: diffuse {r,g,b}; glossy {r,g,b,power}; refractive {ior}; emissive {r,g,b}.
 * normals: per-vertex normals (nv*3) or NULL (flat shading, mesh.h:59-65).  Returns object id. */
uint32_t orc_scene_add_object(void* s, const float* verts, uint32_t nv, const uint32_t* tris,
                              uint32_t nt, const float* normals, int mat, const float* params);
/* Scene::buildKDTree + buildObjKDTrees (scene.cpp:20-65). */
void     orc_scene_build(void* s);
/* Scene::intersect (scene.cpp:102-113), optionally with NoSelfCondition (noself.cpp:13-20).
 * rays: 6 floats per ray; inside/contact: object ids (ORC_NONE = NULL).
 * out: prim (global id), obj, tri, t, b1, b2 (barycentrics of mesh.h:67-72 at o+d*t). */
void     orc_scene_intersect(void* s, const float* rays, const uint32_t* inside,
                             const uint32_t* contact, uint64_t n, int noself, uint32_t* prim,
                             uint32_t* obj, uint32_t* tri, float* t, float* b1, float* b2,
                             int nthreads);
/* Near-ties of closest hits (no Condition): for each ray with a hit (prim, t),
 * the number of other triangles the unpruned traversal (kdtree.cpp:308-347)
 * accepts at |dist - t| < tol; 0 for a miss. */
void     orc_scene_near_ties(void* s, const float* rays, uint64_t n, const uint32_t* prim, const float* t,
                             float tol, uint32_t* count, int nthreads);
/* Scene::findInsideObject (scene.cpp:115-144); cur = ORC_NONE for NULL. */
void     orc_scene_find_inside(void* s, const float* rays, const uint32_t* cur, uint64_t n,
                               uint32_t* out, int nthreads);
/* Mesh::getWorldNormal (mesh.h:47-77) for object obj, local triangle tri, at position p. */
void     orc_world_normal(void* s, uint32_t obj, uint32_t tri, const float* p, float* n_out,
                          float* u1_out, float* u2_out);

/* ---- Path tracer (camera.cpp, mcRenderer.cpp:16-57, pathTracer.cpp:7-76) ---- */
typedef struct {
    float pos[3], focus[3], up[3];
    uint32_t width, height;
    float sight_dist;
} orc_camera;
/* Renders passes [pass0, pass0+npass) of the reference PathTracer estimator with
 * the counter RNG.  For every pixel (row-major, top row first) accumulates, in
 * FP64, sum[3*p+c] += fixVignetting(color) and sumsq[3*p+c] += value^2.
 * The buffers are accumulated into (caller zeroes them). */
void     orc_render(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed,
                    uint32_t pass0, uint32_t npass, double* sum, double* sumsq, int nthreads);
/* Single pixel, single pass: returns the per-sample value and the path length;
 * used by the tests that compare individual paths. */
void     orc_render_sample(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed,
                           uint32_t pixel, uint32_t pass, float* rgb, uint32_t* n_vertices);

/* Camera::generateRay's direction (camera.cpp:13-23) for pixel p and jitter
 * (u, v), and the weight the camera gives the sample: |front . d| (ray.cpp:52-58,
 * camera.cpp:182-186) / |front . d_centre| (fixVignetting, camera.cpp:107-122). */
void     orc_camera_ray(const orc_camera* cam, uint32_t pixel, float u, float v, float* dir, float* weight);
/* AbstractObject::scatter of object self (mcRenderer.cpp:30-31 dispatch) for the
 * ray (o, d) that hits local triangle tri of object obj at distance dist, with
 * insideObj inside, from three uniforms: out = origin, direction, radiance (3
 * each), originProb, directionProb, contact normal (3); ids = insideObj,
 * contactObj.  Returns isDeltaDirection.  diffuseObj.h:34-57, glossyObj.h:39-62,
 * refractiveObj.h:27-88, emissiveObj.h:15-28. */
int      orc_scatter(void* s, uint32_t self, const float* ray6, uint32_t obj, uint32_t tri, float dist,
                     uint32_t inside, const float* rn, float* out, uint32_t* ids);

/* flags of orc_render2 / orc_render_sample2 */
enum {
    ORC_NEE = 1,          /* PathTracer's useNextEventEstimation branch (pathTracer.cpp:39-66) */
    ORC_PATH_VECTOR = 2   /* form the sample from the explicit Path (implied by ORC_NEE) */
};
void     orc_render2(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                     uint32_t npass, uint32_t flags, double* sum, double* sumsq, int nthreads);
void     orc_render_sample2(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed, uint32_t flags,
                            uint32_t pixel, uint32_t pass, float* rgb, uint32_t* n_vertices);
/* Scene::genLightSample (scene.cpp:16-18, surfaceSampler.cpp:17-23,
 * abstractObj.cpp:18-52) from four uniforms (object pick, triangle pick, u1,
 * u2).  Returns 0 when there is nothing to sample (no emissive object, or the
 * picked one has no triangles). */
int      orc_light_sample(void* s, const float rn[4], float* origin, float* normal, float* radiance,
                          float* origin_prob, uint32_t* obj, uint32_t* tri);

/* Light sub-paths (genLightSample + samplePath, bidirPathTracer.cpp:31-33):
 * paths path0 .. path0 + n_paths - 1 of one pass.  rec: 16 words per vertex,
 * max_vertices per path: origin, direction, radiance (3 floats each),
 * originProb, directionProb, contact object, contact triangle (0 without a
 * contact), inside object, isDeltaDirection, 0.  counts: each path's vertex
 * count (at most max_depth + 3; vertices past max_vertices are not stored).
 * Returns the longest path's vertex count. */
uint32_t orc_light_paths(void* s, uint32_t max_depth, uint64_t seed, uint32_t pass, uint32_t path0,
                         uint32_t n_paths, uint32_t max_vertices, float* rec, uint32_t* counts, int nthreads);
/* HashGrid::Reserve(n_cells) + Build(points, radius) (hashgrid.hxx:18-75) and
 * Process (hashgrid.hxx:77-158): per query, the points handed to the query in
 * order (first cap of them stored). */
void*    orc_hashgrid_build(const float* pos, uint64_t n, float radius, uint32_t n_cells);
void     orc_hashgrid_export(void* g, float* bbox, int32_t* cell_ends, int32_t* indices);
void     orc_hashgrid_query(void* g, const float* pos, uint64_t n, const float* q, uint64_t nq, uint32_t cap,
                            uint32_t* counts, int32_t* idx);
void     orc_hashgrid_free(void* g);

/* Motion blur: object obj's mesh (as added) is in its own space and moves
 * with AnimatedTransform(start, t0, end, t1) (motionBlur.cpp; matrices
 * row-major, element (row, col) at [4 * row + col]); the scene shows it at t0
 * until a render, which places every moving object at its time for each pass
 * and rebuilds (Scene::updateSceneForMotionBlur, scene.cpp:67-80). */
void     orc_scene_set_motion(void* s, uint32_t obj, float t0, float t1, const float* start, const float* end);
void     orc_motion_interpolate(float t0, float t1, const float* start, const float* end, float time, float* out);

/* BidirectionalPathTracer::renderPixels (bidirPathTracer.cpp:10-51): per pass,
 * every pixel's eye path and light path (the light path with the light-path
 * vertex counters), every connection, light-tracing splats; the pass's pixel
 * sums checked (isLegalColor) and vignetting-fixed, then accumulated as in
 * orc_render. */
void     orc_render_bdpt(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, double* sum, double* sumsq, int nthreads);
/* VCMTracer::renderPixels (vcmTracer.cpp:6-84) for passes [pass0, pass0 + npass):
 * per pass the radius orc_vcm_radius(base_radius, pass), one light path per
 * pixel (BDPT's), its light points after width * height default ones, the hash
 * grid over them (width * height cells), then per pixel the vertex connections
 * (BDPT's with vcmMisWeight) and the vertex merges; each pass's pixel values
 * checked, vignetting-fixed and added to sum / sumsq as orc_render_bdpt. */
void     orc_render_vcm(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                        uint32_t npass, float base_radius, double* sum, double* sumsq, int nthreads);
float    orc_vcm_radius(float base_radius, uint32_t pass);
/* PhotonMap::renderPixels (photonMap.cpp:21-148) with PPM: per pass the radius
 * orc_vcm_radius(base_radius, pass), the light points and grid of
 * orc_render_vcm, per pixel the eye merge path (maxPathLen 10) and at its first
 * emitter the emitted radiance or at its first diffuse vertex the kernel
 * density estimate of the photons around it; added to sum / sumsq as
 * orc_render_bdpt. */
void     orc_render_ppm(void* s, const orc_camera* cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                        uint32_t npass, float base_radius, double* sum, double* sumsq, int nthreads);

/* ---- Counter RNG (Philox-4x32-10, Salmon et al., SC'11) ---------------- */
void     orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 4 uniform floats in [0,1) for (seed, pixel, pass, vertex). */
void     orc_rng4(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t vertex, float out[4]);

#ifdef __cplusplus
}
#endif
#endif
