/*
 * oracle.h -- CPU restatement of the RTXplay ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product (librtx.so) never links it.
 *
 * Every function cites the reference file:line it restates.  Citations are
 * relative to the reference repository root (otabuzzman/RTXplay); optx/... is
 * the OptiX program (RTWO), bare names are the CPU tracer (RTOW).
 *
 * How each part is pinned (checks live under tests/):
 *   orc_frand48_*        closed form of POSIX drand48 (glibc erand48 on the same
 *                        48-bit state, rounded to float) + committed KAT.
 *   orc_camera_*         FP64 closed form of optx/camera.h:39-47; disk sampler
 *                        invariants.
 *   orc_sphere_mesh      V=2*4^n+2, F=4*4^n, closed 2-manifold (Euler 2), |v|=r,
 *                        and the dedup known-answer test of optx/reduce.cxx:66-78.
 *   orc_closest_f64_*    brute force in FP64: analytic checks on hand-made
 *                        triangles; BVH traversal == brute force on seeded rays.
 *   orc_watertight_f32   == FP64 brute force away from the edge epsilon; closed
 *                        mesh leaks no ray through shared edges/vertices.
 *   orc_rtow_render      bit-exact against the reference rtow binary built from
 *                        rtow.cxx (oracle/ref.mk) on its own hard-coded image;
 *                        orc_rtow_render_ex runs the same tracer (FP64 on the
 *                        glibc stream is that image), or in FP32, or on
 *                        per-pixel streams over threads.
 *   orc_render           (tests/test_oracle_closed_forms.py) images the
 *                        mathematics fixes, for both program sets: the sky of
 *                        an empty scene with rpp == spp, black (recursive) or
 *                        the throughput (iterative) at the depth limit, the
 *                        cosine-weighted sky mean a(1/6 + 5/6 amb) over a
 *                        Lambertian ground, the mirror direction of a fuzz-0
 *                        metal, the absorbed share of a fuzzy metal, Schlick's
 *                        reflected share with 1/ior on entering a glass slab,
 *                        the guide values of both program sets;
 *                        (tests/test_oracle_vs_rtow.py) statistical agreement
 *                        with orc_rtow_render on rtow.cxx's own scene,
 *                        tessellated, at lp.depth 16 and 3: channel means
 *                        within 3 sigma plus the measured tessellation term,
 *                        25x25 tiles within chi^2 < 1.6 per tile and 5 sigma,
 *                        rays per sample.
 *   orc_scene_rtow       the draw ranges and fixed spheres of optx/rtwo.cxx:
 *                        154-243 (one small sphere of radius .2 per grid cell
 *                        at most, offsets below .9, diffuse albedo in [0,1),
 *                        metal albedo in [.5,1) and fuzz in [0,.5), glass
 *                        1.5), the material mix, the exclusion around
 *                        (4,.2,0); the product's host recipe equals it.
 *   orc_pp_srgb          IEC 61966-2-1 curve evaluated in FP64.
 *   orc_render_frozen    equals orc_render at the scene's own materials (to FP32
 *                        rounding); its albedo dependence is a product (checked
 *                        by exact linearity in one channel) and its FD in fuzz
 *                        is the checker of the adjoint kernel's fuzz gradient.
 */
#ifndef RTX_ORACLE_H
#define RTX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: optx/frand48.h:17-35 ---------------------------------------- */
typedef struct { uint64_t state; } orc_frand48;
void  orc_frand48_init(orc_frand48* r, uint32_t seed);
float orc_frand48_next(orc_frand48* r);
/* k draws for each of n seeds -> out[n*k] (row-major per seed) */
void  orc_frand48_draws(const uint32_t* seeds, uint32_t n, uint32_t k, float* out);

/* ---- Camera: optx/camera.h:30-48 (set), optx/camera.h:67-77 (ray) ------- */
typedef struct {
    float eye[3], pat[3], vup[3];
    float fov, aspratio, aperture, fostance;
    float u[3], v[3], w[3], hvec[3], wvec[3], dvec[3];
} orc_camera;
void orc_camera_set(orc_camera* c, const float eye[3], const float pat[3], const float vup[3],
                    float fov, float aspratio, float aperture, float fostance);
void orc_camera_ray(const orc_camera* c, float s, float t, orc_frand48* rng,
                    float ori[3], float dir[3]);
/* primary rays of sample 0 for every pixel of a w x h image (pix = w*y+x),
 * as __raygen__camera forms them: optx/camera_r.cu:35-73.  rays: 8 floats
 * each (ori, tmin, dir, tmax). */
void orc_primary_rays(const orc_camera* c, uint32_t w, uint32_t h, float* rays);

/* ---- Sphere tessellation: optx/sphere.cxx:28-106 ----------------------- */
/* capacity: verts 3*(2*4^n+2) floats, idx 3*4*4^n uints.  Returns 0 on success. */
int orc_sphere_mesh(float radius, uint32_t ndiv, float* verts, uint32_t* nv,
                    uint32_t* idx, uint32_t* ni);
/* the dedup step alone: optx/sphere.cxx:68-106 (== optx/reduce.cxx:42-64) */
int orc_reduce(const float* vtmp, uint32_t n, float* vces, uint32_t* nv, uint32_t* ices);

/* ---- Scene recipe: optx/rtwo.cxx:137-244, drawn from a Frand48 stream ---- */
enum { ORC_DIFFUSE = 0, ORC_REFLECT = 1, ORC_REFRACT = 2 };
typedef struct {
    int32_t type;
    float   albedo[3];
    float   fuzz;
    float   index;
    float   center[3];
    float   radius;
    int32_t kind;   /* 0 ground, 1 small, 2 big (tessellation level choice) */
} orc_thing;
/* seed: Frand48 seed; half: grid a,b in [-half, half).  Returns count. */
uint32_t orc_scene_rtow(uint32_t seed, int32_t half, orc_thing* out, uint32_t cap);

/* world vertex = R v + t, 3x4 row-major, evaluated left to right without
 * contraction: optx/optics_r.cu:59-75 */
void orc_transform_point(const float t[12], const float v[3], float out[3]);

/* ---- Rays and closest hits ---------------------------------------------- */
typedef struct { float ori[3]; float tmin; float dir[3]; float tmax; } orc_ray;

/* FP64 brute force (Moller-Trumbore in double).  tris: 9 floats per tri.
 * Outputs per ray: t (inf on miss), prim (0xFFFFFFFF on miss), u, v
 * (barycentric weights of vertices 1 and 2), t2 = second-closest distinct hit
 * and decidable (see orc_decidable). nthreads<=0 -> 1. */
void orc_closest_f64_brute(const float* tris, uint32_t ntris, const orc_ray* rays, uint32_t nrays,
                           double* t, uint32_t* prim, double* u, double* v, double* t2,
                           int nthreads);

/* A CPU BVH over a triangle soup (binned SAH, own builder).  The same object
 * serves FP64 closest-hit queries and the FP32 render. */
typedef struct orc_bvh orc_bvh;
orc_bvh* orc_bvh_build(const float* tris, uint32_t ntris);
void     orc_bvh_free(orc_bvh* b);
uint32_t orc_bvh_node_count(const orc_bvh* b);
void orc_closest_f64_bvh(const orc_bvh* b, const orc_ray* rays, uint32_t nrays,
                         double* t, uint32_t* prim, double* u, double* v, double* t2,
                         int nthreads);
/* FP32 watertight test (Woop, Benthin, Wald 2013) through the BVH. */
void orc_closest_f32_bvh(const orc_bvh* b, const orc_ray* rays, uint32_t nrays,
                         float* t, uint32_t* prim, float* u, float* v, int nthreads);
/* decidability rule for bit-exact ID comparison: 1 if min barycentric >= eps,
 * (t2-t1)/t1 > eps and (t1-tmin)/tmin > eps; misses are decidable when t2 is
 * inf, i.e. no hit at all. */
int orc_decidable(double t, double u, double v, double t2, double tmin, double eps);

/* ---- Render with the GPU program semantics (camera_r.cu / optics_r.cu) -- */
typedef struct {
    const float*     tris;      /* world-space, 9 floats per tri */
    const uint32_t*  tri_inst;  /* instance (thing) id per tri */
    uint32_t         ntris;
    const orc_thing* things;    /* material + center per instance */
    uint32_t         nthings;
    const orc_bvh*   bvh;       /* over tris */
    float            ambient[3];/* miss record, optx/rtwo.cxx:110-111 */
    int32_t          iterative; /* 0: optx/camera_r.cu + optics_r.cu; 1: optx/camera_i.cu + optics_i.cu */
} orc_scene;

/* Renders rows [y0,y1) of a w x h image.  Outputs are full-size buffers
 * (only the row range is written): raw 3*w*h, rpp w*h, normals/albedos 3*w*h
 * (may be NULL), sumsq 3*w*h (per-pixel sum of squared sample colours, may be
 * NULL).  Pixel streams are seeded with pix + seed_base (0 = the reference's
 * seeding; another base gives independent streams, e.g. for a high-spp
 * reference image).  Returns total rays traced in the range. */
uint64_t orc_render(const orc_scene* s, const orc_camera* cam, uint32_t w, uint32_t h,
                    uint32_t spp, uint32_t depth, uint32_t y0, uint32_t y1, int nthreads,
                    float* raw, uint32_t* rpp, float* normals, float* albedos, float* sumsq,
                    uint32_t seed_base);

/* Frozen-structure render for finite differences of material parameters:
 * each sample path is traced as orc_render traces it (the scene's own
 * materials decide the path), then its radiance is evaluated in FP64 along
 * the recorded vertices with the materials `params` (float4 per thing:
 * albedo rgb, fuzz), without re-tracing: hits, normals, random vectors and
 * discrete choices stay fixed.  With params equal to the scene's materials
 * this is orc_render's image up to FP32-vs-FP64 rounding.  raw: 3*w*h
 * clamped means (rows [y0,y1) written).  depth <= 64. */
void orc_render_frozen(const orc_scene* s, const orc_camera* cam, uint32_t w, uint32_t h, uint32_t spp,
                       uint32_t depth, uint32_t y0, uint32_t y1, int nthreads, const float* params, double* raw);

/* ---- Restatement of rtow.cxx (FP64, analytic spheres, glibc rand) -------- */
/* argorder: 0 = function arguments evaluated left to right, 1 = right to left
 * (the reference leaves the order to the compiler).  Writes w*h*3 bytes in
 * PPM order (top row first).  Returns rays traced.  Calls srand(1) first. */
uint64_t orc_rtow_render(int w, int h, int spp, int depth, int argorder, uint8_t* rgb);
/* rays traced for the first npix pixels of that image (its output order) */
uint64_t orc_rtow_rays(int w, int h, int spp, int depth, int argorder, uint64_t npix);
/* the rtow scene (rtow.cxx:51-80) as things, same argorder convention */
uint32_t orc_rtow_scene(int argorder, orc_thing* out, uint32_t cap);
/* The same tracer with options.  fp32: 0 = FP64 (the reference's double), 1 =
 * every operation in float.  nthreads: 0 = one thread drawing from glibc
 * rand() after srand(1) in the reference's pixel order (fp32 = 0 gives the
 * reference's image byte for byte); >= 1 = that many threads, each pixel
 * drawing from its own drand48 stream seeded with pix + seed_base (pix =
 * w*y + x, y = 0 the bottom row).  Pixel mapping, camera, scene and depth
 * semantics are rtow.cxx's: s = 2(x+u)/(w-1)-1, trace(ray, depth) allows
 * depth scatters (depth+1 rays).  Outputs (any may be NULL): rgb w*h*3
 * bytes in PPM order (top row first, gamma 2), mean 3*w*h doubles (the
 * pixel's colour/spp before gamma, pix order), sumsq 3*w*h (sum over
 * samples of the squared sample colour), rpp w*h (rays per pixel).
 * Returns the rays traced. */
typedef struct { int fp32; int nthreads; int argorder; uint32_t seed_base; } orc_rtow_opts;
uint64_t orc_rtow_render_ex(const orc_rtow_opts* o, int w, int h, int spp, int depth, uint8_t* rgb, double* mean,
                            double* sumsq, uint32_t* rpp);

/* ---- Scatter building blocks used by orc_render (for known-answer tests) */
void  orc_reflect(const float d[3], const float n[3], float out[3]);              /* of unit(d) */
void  orc_refract(const float d[3], const float n[3], float ratio, float out[3]); /* of unit(d) */
float orc_schlick(float cos_t, float ratio);
void  orc_sky(const float d[3], const float ambient[3], float out[3]);

/* ---- Post-process: optx/postproc.cu:1-61 -------------------------------- */
void orc_pp_srgb(const float* src, uint8_t* dst, int w, int h);      /* FP32 like the kernel */
void orc_pp_srgb_f64(const float* src, double* dst, int w, int h);    /* exact curve, x255 */
void orc_pp_none(const float* src, uint8_t* dst, int w, int h);

#ifdef __cplusplus
}
#endif
#endif
