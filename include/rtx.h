/*
 * rtx.h -- C-ABI of the MI355X ray-tracing hot path (librtx.so).
 *
 * Plain C types only: opaque handles, pointers and sizes.  Device pointers are
 * HIP device addresses; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every call returns an rtx_status; on failure rtx_last_error()
 * returns a message for the calling thread.
 *
 * Each entry point names the reference interface it replaces.  Citations are
 * relative to the reference repository (otabuzzman/RTXplay); optx/... is the
 * OptiX program.
 */
#ifndef RTX_H
#define RTX_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------- */
typedef enum {
    RTX_SUCCESS = 0,
    RTX_ERROR_INVALID_VALUE = 1,   /* bad argument (null pointer, size 0, spp 0, ...) */
    RTX_ERROR_INVALID_HANDLE = 2,  /* unknown traversable handle / id out of range */
    RTX_ERROR_NOT_BUILT = 3,       /* scene has no acceleration structure yet */
    RTX_ERROR_OUT_OF_MEMORY = 4,
    RTX_ERROR_HIP = 5,             /* a HIP runtime call failed; message has the text */
    RTX_ERROR_LIMIT = 6            /* a documented size limit is exceeded */
} rtx_status;

/* message of the last failed call on this thread ("" if none) */
const char* rtx_last_error(void);

/* ---- value types, layout-compatible with the reference's CUDA structs ----- */
typedef struct { float x, y, z; } rtx_float3;           /* CUDA float3 */
typedef struct { uint32_t x, y, z; } rtx_uint3;         /* CUDA uint3  */
typedef struct { uint8_t x, y, z, w; } rtx_uchar4;      /* CUDA uchar4 */

/* Optics / Thing: optx/thing.h:17-52 (same field order, types and union) */
enum { RTX_OPTICS_DIFFUSE = 0, RTX_OPTICS_REFLECT = 1, RTX_OPTICS_REFRACT = 2, RTX_OPTICS_NUM = 3 };
typedef struct { rtx_float3 albedo; } rtx_diffuse;
typedef struct { rtx_float3 albedo; float fuzz; } rtx_reflect;
typedef struct { float index; } rtx_refract;
typedef struct {
    int32_t type;
    union { rtx_diffuse diffuse; rtx_reflect reflect; rtx_refract refract; };
} rtx_optics;
typedef struct {
    rtx_float3* vces;   /* non-owning device pointer, filled by rtx_thing_add */
    rtx_uint3*  ices;   /* non-owning device pointer, filled by rtx_thing_add */
    rtx_optics  optics;
} rtx_thing;

/* Camera: optx/camera.h:81-93 (private members of the class, as a POD) */
typedef struct {
    rtx_float3 eye, pat, vup;
    float fov, aspratio, aperture, fostance;
    rtx_float3 u, v, w;
    rtx_float3 hvec, wvec, dvec;
} rtx_camera;

/* Launch parameters: optx/rtwo.h:26-52, same field order and types with
 * OptixTraversableHandle -> uint64_t (a handle from rtx_scene_build). */
typedef struct {
    rtx_uchar4* image;          /* not written by rtx_launch (postproc output) */
    uint32_t    image_w;
    uint32_t    image_h;
    rtx_float3* rawRGB;         /* w*h mean colour, clamped to [0,1] */
    rtx_float3* normals;        /* optional (NULL: not written) */
    rtx_float3* albedos;        /* optional (NULL: not written) */
    rtx_camera  camera;
    uint32_t    spp;
    uint32_t    depth;          /* max rays per sample (>=1) */
    uint64_t    is_handle;
    uint32_t*   rpp;            /* rays per pixel; optional (NULL: not written) */
    bool        picker;
    uint32_t    pick_x;
    uint32_t    pick_y;
    uint32_t*   pick_id;
} rtx_lp_general;

/* A ray for the closest-hit query; 32 bytes. */
typedef struct { rtx_float3 ori; float tmin; rtx_float3 dir; float tmax; } rtx_ray;
/* Closest hit: t, barycentrics (u,v) weighting vertices 1 and 2 (as
 * optixGetTriangleBarycentrics), instance id and primitive index within the
 * instance's mesh (as optixGetInstanceId / optixGetPrimitiveIndex).  On a
 * miss instance_id == primitive_index == 0xFFFFFFFF and t = +inf. */
typedef struct { float t, u, v; uint32_t instance_id; uint32_t primitive_index; uint32_t prim; uint32_t pad[2]; } rtx_hit;

typedef struct rtx_ctx rtx_ctx;

/* ---- context ------------------------------------------------------------- */
/* Replaces optixInit + optixDeviceContextCreate (optx/rtwo.cxx:115-128) and
 * owns what Scene owns in the reference (optx/scene.h:37-62). */
rtx_status rtx_ctx_create(int device, rtx_ctx** ctx);
rtx_status rtx_ctx_destroy(rtx_ctx* ctx);

/* ---- camera ------------------------------------------------------------- */
/* Camera::set, optx/camera.h:30-48 (host side, FP32) */
rtx_status rtx_camera_set(rtx_camera* cam, rtx_float3 eye, rtx_float3 pat, rtx_float3 vup,
                          float fov, float aspratio, float aperture, float fostance);

/* ---- geometry ------------------------------------------------------------ */
/* Sphere(radius, ndiv).mesh(), optx/sphere.cxx:17-106 (host).  With vces/ices
 * NULL only the counts are returned.  Capacities in elements. */
rtx_status rtx_sphere_mesh(float radius, uint32_t ndiv, rtx_float3* vces, uint32_t vcap,
                           uint32_t* nv, rtx_uint3* ices, uint32_t icap, uint32_t* ni);

/* Object(wavefront), optx/object.cxx:23-93 (host): reads a Wavefront OBJ file
 * into one indexed triangle mesh per shape ("o"/"g"), each shape's vertices
 * in first-use order and its indices renumbered into them (the reference's
 * index-bound and renumbering defects are not reproduced).  Faces of n > 3
 * corners become the fan (0, k, k+1).  rtx_object_mesh returns pointers into
 * the object, valid until rtx_object_free. */
typedef struct rtx_object rtx_object;
rtx_status rtx_object_load(const char* wavefront, rtx_object** obj);
rtx_status rtx_object_size(const rtx_object* obj, uint32_t* n);
rtx_status rtx_object_mesh(const rtx_object* obj, uint32_t m, const rtx_float3** vces, uint32_t* nv,
                           const rtx_uint3** ices, uint32_t* ni);
rtx_status rtx_object_free(rtx_object* obj);
/* Writes meshes as one OBJ shape named `name`: all vertices, then all faces
 * (1-based, offset per mesh), with vertex and triangle count comments, as
 * the sphere and object writers of optx/sphere.cxx:108-140 and
 * optx/object.cxx:96-122 print them (coordinates with 9 significant digits,
 * so a float reads back exactly). */
rtx_status rtx_object_write(const char* path, const char* name, const rtx_float3* const* vces, const uint32_t* nv,
                            const rtx_uint3* const* ices, const uint32_t* ni, uint32_t nmesh);

/* Scene::add(Object&), optx/scene.cxx:38-167: uploads one mesh (the caller
 * concatenates an Object's submeshes) and returns its id. */
rtx_status rtx_mesh_add(rtx_ctx* ctx, const rtx_float3* vces, uint32_t nv,
                        const rtx_uint3* ices, uint32_t ni, uint32_t* mesh_id);

/* Scene::add(Thing&, object), optx/scene.cxx:169-196: new instance of a mesh
 * with identity transform; fills thing->vces/ices with the mesh's device
 * buffers.  Returns the instance (thing) id. */
rtx_status rtx_thing_add(rtx_ctx* ctx, rtx_thing* thing, uint32_t mesh_id, uint32_t* thing_id);

/* Scene::set / Scene::get, optx/scene.cxx:198-225: 3x4 row-major object-to-
 * world transform.  An id out of range returns RTX_ERROR_INVALID_HANDLE (the
 * reference returns false). */
rtx_status rtx_thing_set(rtx_ctx* ctx, uint32_t thing_id, const float transform[12]);
rtx_status rtx_thing_get(rtx_ctx* ctx, uint32_t thing_id, float transform[12]);
/* Scene::operator[], optx/scene.h:21 */
rtx_status rtx_thing_at(rtx_ctx* ctx, uint32_t thing_id, rtx_thing* thing);
rtx_status rtx_thing_count(rtx_ctx* ctx, uint32_t* n);

/* Layout of the acceleration structure the next rtx_scene_build makes.  The
 * reference always builds an IAS over one GAS per mesh (optx/scene.cxx:38-273);
 * RTX_BUILD_TWO_LEVEL is that design: one BVH per mesh over its object-space
 * triangles, one over the instances' world boxes, rays transformed into an
 * instance's object space on entry, and rtx_scene_update a refit of the top
 * level.  RTX_BUILD_FLAT (the default) transforms every instance's triangles
 * to world space and builds one BVH over all of them: more memory for
 * instanced meshes, fewer node visits per ray, and rtx_scene_update a
 * rebuild.  Both give the same hits up to float rounding. */
enum { RTX_BUILD_FLAT = 0, RTX_BUILD_TWO_LEVEL = 1 };
rtx_status rtx_scene_set_build_mode(rtx_ctx* ctx, int mode);
rtx_status rtx_scene_get_build_mode(rtx_ctx* ctx, int* mode);

/* Scene::build, optx/scene.cxx:227-273: builds the acceleration structure on
 * the device (binned SAH BVH, layout by the build mode).  Synchronous.
 * Returns the traversable handle. */
rtx_status rtx_scene_build(rtx_ctx* ctx, void* stream, uint64_t* is_handle);
/* Scene::update, optx/scene.cxx:275-294: brings the acceleration structure up
 * to date with rtx_thing_set changes and keeps the handle.  A two-level scene
 * whose instance count is unchanged refits its top level in place (new
 * instance boxes, same topology, as OPTIX_BUILD_OPERATION_UPDATE); otherwise
 * the scene is rebuilt.  Material records return to the things' values.
 * Synchronous. */
rtx_status rtx_scene_update(rtx_ctx* ctx, uint64_t is_handle, void* stream);

/* Statistics of a built scene (for reports and tests).  num_prims counts
 * the triangles of all instances; for a two-level scene the node and leaf
 * counts add up both levels, max_depth is top + deepest bottom, sah_cost
 * is the top level's. */
typedef struct {
    uint32_t num_prims, num_nodes, num_leaves, max_depth, max_leaf;
    float    sah_cost;          /* SAH cost, C_trav = 1, C_isect = 1 */
    float    build_ms;          /* device time of the last build */
    uint64_t device_bytes;      /* nodes + leaf-ordered triangles + instance data */
    uint32_t wide_nodes;        /* 4-wide nodes the traversal uses (collapsed from the binary tree) */
    uint32_t wide_depth;        /* levels of the 4-wide tree (two-level: top + deepest bottom) */
    uint32_t two_level;         /* 1: built with RTX_BUILD_TWO_LEVEL */
    uint32_t unique_prims;      /* triangles stored (two-level: of the meshes in use, once each) */
    uint32_t tlas_nodes;        /* two-level: 4-wide nodes of the top level */
    uint32_t num_blas;          /* two-level: bottom-level trees (meshes in use) */
} rtx_scene_stats;
rtx_status rtx_scene_stats_get(rtx_ctx* ctx, uint64_t is_handle, rtx_scene_stats* st);

/* Copies the world-space triangles (9 floats each, original primitive order:
 * instance-major, mesh triangle order) and per-triangle instance ids to host
 * memory.  Either pointer may be NULL. */
rtx_status rtx_scene_triangles(rtx_ctx* ctx, uint64_t is_handle, float* tris, uint32_t* tri_inst);

/* Launch tuning of a context (no reference counterpart).  shade_batch: in
 * the render kernel a wave traverses until at least this many of its lanes
 * hold a final closest hit, shades those and resumes the others (1..64; 64 =
 * every lane of the wave finishes first; default 40).  inner_frac: a
 * traversal round's node phase ends when fewer than inner_frac/64 of the
 * traversing lanes can still step through inner nodes (each lane parks up
 * to two leaves and goes on); the parked leaves are then tested (1..64; 1 = every lane
 * steps as far as it can first, 64 = the first lane that must wait ends the
 * phase; default 36).  tile_w: pixels are handed
 * out in tiles of tile_w x (64 / tile_w) (1, 2, 4, ..., 64; default 8).
 * stack_lds: traversal stack entries per lane kept in LDS, the rest in a
 * global overflow area (16 or 32; default 32).  blocks_per_cu: persistent
 * grid of blocks_per_cu x CUs blocks of 256 lanes (1..max_blocks_per_cu, the
 * occupancy limit, which is also the default; max_blocks_per_cu is read
 * only).  0 fields keep the current value. */
typedef struct {
    uint32_t shade_batch;
    uint32_t inner_frac;
    uint32_t tile_w;
    uint32_t stack_lds;
    uint32_t blocks_per_cu;
    uint32_t max_blocks_per_cu;
    uint32_t reserved[2];
} rtx_tuning;
rtx_status rtx_tuning_get(rtx_ctx* ctx, rtx_tuning* t);
rtx_status rtx_tuning_set(rtx_ctx* ctx, const rtx_tuning* t);

/* Miss record (SBT miss data), optx/rtwo.cxx:110-111 and :474-489.  Default {.5,.7,1}. */
rtx_status rtx_set_miss_color(rtx_ctx* ctx, rtx_float3 ambient);

/* Which of the reference's two program sets a context's launches follow
 * (optx/rtwo.cxx:44-50 picks one at compile time; optx/Makefile:57-59).
 * RTX_SEMANTICS_RECURSIVE (the default here) is the -DRECURSIVE build
 * (optx/camera_r.cu, optx/optics_r.cu): a hit by the lp.depth-th ray of a
 * sample is black and draws nothing (optx/optics_r.cu:31-42); the guide
 * values are the deepest diffuse or metal hit's normal and the deepest
 * non-zero albedo (set after the recursion returns, optx/optics_r.cu:143-146).
 * RTX_SEMANTICS_ITERATIVE is the default build (optx/camera_i.cu,
 * optx/optics_i.cu): that hit is shaded like any other (its random draws
 * included) and the sample keeps the throughput as its colour
 * (optx/camera_i.cu:75-95); the guide values are the first diffuse or metal
 * hit's normal and the first non-zero albedo (optx/optics_i.cu:98-101).
 * Misses, ray counts and every other step are the same.  Applies to
 * rtx_launch, rtx_launch_counted, rtx_launch_adjoint and batches launched
 * after the call. */
enum { RTX_SEMANTICS_RECURSIVE = 0, RTX_SEMANTICS_ITERATIVE = 1 };
rtx_status rtx_set_semantics(rtx_ctx* ctx, int mode);
rtx_status rtx_get_semantics(rtx_ctx* ctx, int* mode);

/* ---- launch -------------------------------------------------------------- */
/* Launcher::ignite, optx/launcher.cxx:49-77 (optixLaunch of the raygen program
 * optx/camera_r.cu:23-121 with the closest-hit programs optx/optics_r.cu).
 * Asynchronous on `stream` (the C++ Launcher syncs, as the reference does).
 * once != 0 launches the 1x1 picker path (lp->picker).  A frame of 16 or more
 * spp and at least two pixels per lane of the persistent grid first runs a
 * one-sample probe and renders its 64-pixel tiles in descending order of the
 * probe's rays (the pixels do not depend on the order); the first such launch
 * of a context allocates the order's arrays (8 bytes per tile). */
rtx_status rtx_launch(rtx_ctx* ctx, const rtx_lp_general* lp, int once, void* stream);

/* rtx_launch without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same pixels as rtx_launch; every scratch buffer of the
 * launch (work counter, probe counts and tile order, traversal-stack
 * overflow) lives in one static device array of 48 MiB per GPU, which the
 * library holds until the process ends.  Limits, each rejected with
 * RTX_ERROR_LIMIT before anything is launched: a frame rendered in probe
 * order (as rtx_launch describes) may have at most 2^18 tiles (2^24
 * pixels), and the scene's stack overflow (entries beyond the context's
 * stack_lds, times the persistent grid's lanes, 4 bytes each) must fit the
 * 46 MiB stack part.  All calls on one GPU share that GPU's array, which
 * keeps the last call's contents (the launch clears what must start at zero
 * with its own kernels): calls that overlap on one GPU race. */
rtx_status rtx_launch_noalloc(rtx_ctx* ctx, const rtx_lp_general* lp, int once, void* stream);

/* Batched render (no reference counterpart; for callers that render many
 * small frames, e.g. 64 uneven scenes at once): n frames, each
 * with its own launch parameters (scene handle, camera, size, spp, depth,
 * output buffers), rendered by ONE persistent kernel launch that pulls
 * 64-pixel tiles of all frames from one work counter, largest frames first.
 * Every frame's pixels are identical to rtx_launch of the same parameters.
 * rtx_batch_create validates the frames and uploads the job table (the
 * context's tuning at that moment applies); with use_graph != 0 the first
 * rtx_batch_launch on a stream captures the launch into a hipGraph, which
 * later launches replay (a graph needs a non-NULL stream; with NULL the
 * batch launches directly).  Launches of one batch, and launches of its
 * context, must not overlap. */
typedef struct rtx_batch rtx_batch;
rtx_status rtx_batch_create(rtx_ctx* ctx, const rtx_lp_general* lps, uint32_t n, int use_graph, rtx_batch** batch);
rtx_status rtx_batch_launch(rtx_batch* batch, void* stream);
rtx_status rtx_batch_destroy(rtx_batch* batch);

/* ---- differentiable rendering (no reference counterpart) ------------------
 * Material parameters are float4 per instance (instance = thing id):
 * (albedo r, g, b, fuzz).  rtx_scene_set_materials writes them into the
 * built scene on the device (diffuse instances take the albedo, metal
 * instances albedo and fuzz, dielectric instances nothing); they hold until
 * the next rtx_scene_build / rtx_scene_update.  rtx_scene_get_materials
 * reads the built scene's values (for a dielectric: 0, 0, 0, ior).  params
 * are device pointers to num_things x 4 floats. */
rtx_status rtx_scene_set_materials(rtx_ctx* ctx, uint64_t is_handle, const float* params, void* stream);
rtx_status rtx_scene_get_materials(rtx_ctx* ctx, uint64_t is_handle, float* params, void* stream);

/* Adjoint of rtx_launch w.r.t. the material parameters (path replay): given
 * the forward image in lp->rawRGB (the output of rtx_launch with the same
 * lp) and dL_drgb = dL/d rawRGB (device float3 per pixel), writes
 * d_params = dL/d params (device float4 per instance: d albedo rgb, d fuzz;
 * overwritten).  The kernel replays every sample path with the forward's
 * per-pixel random streams and propagates the loss gradient back along each
 * path that reaches the sky: albedo by the product rule, fuzz through the
 * chain of metal and dielectric directions to the sky colour.  Gradients
 * hold the discrete path structure fixed (hit triangles, reflect/refract and
 * absorb decisions), the interior derivative; pixels whose mean clamps to 0
 * or 1 pass no gradient.  Sums use float atomics, so the last bits vary
 * from run to run.  Frames that rtx_launch renders in probe order are
 * replayed in the same kind of order. */
rtx_status rtx_launch_adjoint(rtx_ctx* ctx, const rtx_lp_general* lp, const rtx_float3* dL_drgb, float* d_params,
                              void* stream);

/* Statistics launch: the same render as rtx_launch, which also writes per
 * pixel the number of BVH node visits and triangle tests of all its rays
 * (device uint64 arrays of image_w*image_h) and, if wave_cycles is not NULL,
 * adds to wave_cycles[0] / [1] the shader-clock cycles the waves spent in
 * traversal / in the whole render loop and to [2] / [3] the wave-level
 * iterations of the inner-node / triangle loops (4 device uint64, zeroed by
 * the caller).  No
 * counterpart in the reference; it feeds the bytes-per-ray model of the
 * benchmark. */
rtx_status rtx_launch_counted(rtx_ctx* ctx, const rtx_lp_general* lp, uint64_t* node_visits,
                              uint64_t* tri_tests, uint64_t* wave_cycles, void* stream);

/* Closest-hit query for a batch of rays (the optixTrace call the programs
 * make, optx/camera_r.cu:82-99), for tests and hybrid callers.  rays and hits
 * are device pointers.  Asynchronous. */
rtx_status rtx_trace(rtx_ctx* ctx, uint64_t is_handle, const rtx_ray* rays, uint32_t n,
                     rtx_hit* hits, void* stream);

/* Device Frand48 (optx/frand48.h:17-35): k draws after init(seed) for each of
 * n seeds into out[n*k] (device pointers).  Asynchronous. */
rtx_status rtx_rng_draws(const uint32_t* seeds, uint32_t n, uint32_t k, float* out, void* stream);

/* ---- post-process -------------------------------------------------------- */
/* pp_sRGB / pp_none, optx/postproc.cu:49-61 (the reference launches on the null
 * stream and reports nothing; these take a stream and return a status). */
rtx_status rtx_pp_srgb(const rtx_float3* src, rtx_uchar4* dst, int w, int h, void* stream);
rtx_status rtx_pp_none(const rtx_float3* src, rtx_uchar4* dst, int w, int h, void* stream);

/* ---- device memory plumbing for callers without a HIP binding ------------ */
rtx_status rtx_malloc(void** dptr, size_t bytes);
rtx_status rtx_free(void* dptr);
rtx_status rtx_memcpy_htod(void* dst, const void* src, size_t bytes);
rtx_status rtx_memcpy_dtoh(void* dst, const void* src, size_t bytes);
rtx_status rtx_memset(void* dptr, int value, size_t bytes);
rtx_status rtx_synchronize(void);
rtx_status rtx_device_count(int* n);

#ifdef __cplusplus
}
#endif
#endif /* RTX_H */
