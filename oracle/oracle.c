/*
 * oracle.c -- CPU restatement of the RTXplay hot path (test infrastructure).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this code, as the checker.  See oracle.h for how each part is pinned.
 * Build: oracle/Makefile (gcc, -ffp-contract=off so every float expression is
 * evaluated exactly as written, one rounding per operation).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ========================================================================= */
/* Frand48: optx/frand48.h:17-35                                              */
/* ========================================================================= */

float orc_frand48_next(orc_frand48* r) {
    /* optx/frand48.h:27-31: 64-bit LCG with the drand48 constants; the low
     * 48 bits, converted to float (round to nearest), scaled by 2^-48. */
    r->state = 0x5DEECE66Dull * r->state + 0xBull;
    return (float)(r->state & 0xFFFFFFFFFFFFull) / (float)(0xFFFFFFFFFFFFull + 1ull);
}

void orc_frand48_init(orc_frand48* r, uint32_t seed) {
    /* optx/frand48.h:19-25: state = seed, then 16 discarded draws */
    r->state = seed;
    for (int i = 0; i < 16; i++) (void)orc_frand48_next(r);
}

void orc_frand48_draws(const uint32_t* seeds, uint32_t n, uint32_t k, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        orc_frand48 r;
        orc_frand48_init(&r, seeds[i]);
        for (uint32_t j = 0; j < k; j++) out[(size_t)i * k + j] = orc_frand48_next(&r);
    }
}

/* util::rnd(min,max): optx/util_gpu.h:28 */
static inline float rnd_range(orc_frand48* r, float lo, float hi) {
    return lo + orc_frand48_next(r) * (hi - lo);
}

/* ========================================================================= */
/* float3 helpers with the V:: semantics of optx/v.h:26-53                    */
/* ========================================================================= */

typedef struct { float x, y, z; } f3;

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 scl(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len(f3 a) { return sqrtf(dot(a, a)); }
static inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* optx/v.h:51: 1/len * v */
static inline f3 unitv(f3 a) { return scl(1.f / len(a), a); }
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static inline void st3(float* p, f3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline int isnull3(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }

/* optx/v.h:72 -- rejection sampling in the cube [-1,1]^3, draws x then y then z */
static f3 rnd_in_sphere(orc_frand48* r) {
    for (;;) {
        float x = rnd_range(r, -1.f, 1.f);
        float y = rnd_range(r, -1.f, 1.f);
        float z = rnd_range(r, -1.f, 1.f);
        f3 v = mk(x, y, z);
        if (1.f > dot(v, v)) return v;
    }
}
/* optx/v.h:73 */
static f3 rnd_on_sphere(orc_frand48* r) { return unitv(rnd_in_sphere(r)); }
/* optx/v.h:75 -- rejection sampling in the square, draws x then y */
static f3 rnd_in_disk(orc_frand48* r) {
    for (;;) {
        float x = rnd_range(r, -1.f, 1.f);
        float y = rnd_range(r, -1.f, 1.f);
        f3 v = mk(x, y, 0.f);
        if (1.f > dot(v, v)) return v;
    }
}

/* ========================================================================= */
/* Camera: optx/camera.h:30-48 and 67-77                                     */
/* ========================================================================= */

void orc_camera_set(orc_camera* c, const float eye[3], const float pat[3], const float vup[3],
                    float fov, float aspratio, float aperture, float fostance) {
    memcpy(c->eye, eye, 12); memcpy(c->pat, pat, 12); memcpy(c->vup, vup, 12);
    c->fov = fov; c->aspratio = aspratio; c->aperture = aperture; c->fostance = fostance;
    f3 E = ld3(eye), P = ld3(pat), U = ld3(vup);
    f3 w = unitv(sub(E, P));                       /* optx/camera.h:39 */
    f3 u = unitv(cross(U, w));                     /* optx/camera.h:40 */
    f3 v = cross(w, u);                            /* optx/camera.h:41 */
    const float kPi = 3.14159265358979323846f;     /* optx/util.h:22 */
    float rad = fov * kPi / 180.f;                 /* optx/util_gpu.h:31 */
    float h = 2.f * tanf(.5f * rad);               /* optx/camera.h:43 */
    float ww = h * aspratio;                       /* optx/camera.h:44 */
    st3(c->u, u); st3(c->v, v); st3(c->w, w);
    st3(c->hvec, scl(fostance * h / 2.f, v));      /* optx/camera.h:45 */
    st3(c->wvec, scl(fostance * ww / 2.f, u));     /* optx/camera.h:46 */
    st3(c->dvec, scl(fostance, w));                /* optx/camera.h:47 */
}

void orc_camera_ray(const orc_camera* c, float s, float t, orc_frand48* rng,
                    float ori[3], float dir[3]) {
    f3 r = scl(c->aperture / 2.f, rnd_in_disk(rng));                 /* optx/camera.h:72 */
    f3 o = add(scl(r.x, ld3(c->u)), scl(r.y, ld3(c->v)));            /* optx/camera.h:73 */
    st3(ori, add(ld3(c->eye), o));                                   /* optx/camera.h:75 */
    f3 d = sub(sub(add(scl(s, ld3(c->wvec)), scl(t, ld3(c->hvec))), ld3(c->dvec)), o);
    st3(dir, d);                                                     /* optx/camera.h:76 */
}

/* optx/camera_r.cu:69-70: s/t from the pixel and two jitter draws */
static inline void pixel_st(uint32_t x, uint32_t y, uint32_t w, uint32_t h, orc_frand48* rng,
                            float* s, float* t) {
    *s = 2.f * ((float)x + orc_frand48_next(rng)) / (float)w - 1.f;
    *t = 2.f * ((float)y + orc_frand48_next(rng)) / (float)h - 1.f;
}

void orc_primary_rays(const orc_camera* c, uint32_t w, uint32_t h, float* rays) {
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            uint32_t pix = w * y + x;
            orc_frand48 rng;
            orc_frand48_init(&rng, pix);
            float s, t;
            pixel_st(x, y, w, h, &rng, &s, &t);
            float* r = rays + (size_t)pix * 8;
            orc_camera_ray(c, s, t, &rng, r, r + 4);
            r[3] = 1e-3f;
            r[7] = 1e16f;
        }
}

/* ========================================================================= */
/* Sphere tessellation: optx/sphere.cxx:28-106                                */
/* ========================================================================= */

typedef struct { float* v; size_t n, cap; } vbuf;

static void vpush(vbuf* b, f3 a) {
    if (b->n + 3 > b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->v = (float*)realloc(b->v, b->cap * sizeof(float));
    }
    b->v[b->n++] = a.x; b->v[b->n++] = a.y; b->v[b->n++] = a.z;
}

/* optx/sphere.cxx:51-66 */
static void pumpup(vbuf* out, float radius, f3 a, f3 b, f3 c, uint32_t cdiv) {
    if (cdiv > 0) {
        f3 ab = unitv(scl(.5f, add(a, b)));
        f3 bc = unitv(scl(.5f, add(b, c)));
        f3 ca = unitv(scl(.5f, add(c, a)));
        pumpup(out, radius, a, ab, ca, cdiv - 1);
        pumpup(out, radius, ab, b, bc, cdiv - 1);
        pumpup(out, radius, ca, bc, c, cdiv - 1);
        pumpup(out, radius, ab, bc, ca, cdiv - 1);
    } else {
        vpush(out, scl(radius, a));
        vpush(out, scl(radius, b));
        vpush(out, scl(radius, c));
    }
}

/* Exact float3 dedup in first-seen order (optx/sphere.cxx:68-106).  The
 * reference uses a std::set ordered lexicographically on (x,y,z); an open
 * addressing hash on the bit patterns gives the same first-seen numbering.
 * +0.f and -0.f compare equal there, so the key canonicalises -0 to +0. */
static inline uint32_t fbits(float f) { if (f == 0.f) f = 0.f; uint32_t u; memcpy(&u, &f, 4); return u; }

int orc_reduce(const float* vtmp, uint32_t n, float* vces, uint32_t* nv, uint32_t* ices) {
    uint32_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    uint32_t* slot = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    if (!slot) return -1;
    memset(slot, 0xFF, sizeof(uint32_t) * cap);
    uint32_t count = 0;
    for (uint32_t i = 0; i < n; i++) {
        const float* p = vtmp + 3 * (size_t)i;
        uint32_t kx = fbits(p[0]), ky = fbits(p[1]), kz = fbits(p[2]);
        uint32_t hsh = (kx * 73856093u) ^ (ky * 19349663u) ^ (kz * 83492791u);
        uint32_t s = hsh & (cap - 1);
        for (;;) {
            uint32_t id = slot[s];
            if (id == 0xFFFFFFFFu) {
                slot[s] = count;
                vces[3 * count + 0] = p[0]; vces[3 * count + 1] = p[1]; vces[3 * count + 2] = p[2];
                ices[i] = count++;
                break;
            }
            const float* q = vces + 3 * (size_t)id;
            if (q[0] == p[0] && q[1] == p[1] && q[2] == p[2]) { ices[i] = id; break; }
            s = (s + 1) & (cap - 1);
        }
    }
    *nv = count;
    free(slot);
    return 0;
}

int orc_sphere_mesh(float radius, uint32_t ndiv, float* verts, uint32_t* nv,
                    uint32_t* idx, uint32_t* ni) {
    const float m = .57735026919f;                 /* optx/sphere.cxx:34 */
    f3 v00 = mk(m, m, m), v01 = mk(m, -m, -m), v02 = mk(-m, -m, m), v03 = mk(-m, m, -m);
    vbuf b = {0, 0, 0};
    pumpup(&b, radius, v00, v01, v02, ndiv);       /* optx/sphere.cxx:43-46 */
    pumpup(&b, radius, v00, v02, v03, ndiv);
    pumpup(&b, radius, v00, v03, v01, ndiv);
    pumpup(&b, radius, v03, v02, v01, ndiv);
    uint32_t n = (uint32_t)(b.n / 3);
    int rc = orc_reduce(b.v, n, verts, nv, idx);
    *ni = n / 3;
    free(b.v);
    return rc;
}

/* ========================================================================= */
/* Scene recipe: optx/rtwo.cxx:137-244, drawn from one Frand48 stream.         */
/* Draw order (fixed here; the reference's albedo products leave it to the    */
/* compiler): select, center x, center z, then the material draws.            */
/* ========================================================================= */

static void set_thing(orc_thing* t, int type, f3 albedo, float fuzz, float index, f3 c,
                      float radius, int kind) {
    memset(t, 0, sizeof(*t));
    t->type = type;
    t->albedo[0] = albedo.x; t->albedo[1] = albedo.y; t->albedo[2] = albedo.z;
    t->fuzz = fuzz; t->index = index;
    t->center[0] = c.x; t->center[1] = c.y; t->center[2] = c.z;
    t->radius = radius; t->kind = kind;
}

uint32_t orc_scene_rtow(uint32_t seed, int32_t half, orc_thing* out, uint32_t cap) {
    orc_frand48 rng;
    orc_frand48_init(&rng, seed);
    uint32_t n = 0;
#define PUSH(...) do { if (n < cap) set_thing(&out[n], __VA_ARGS__); n++; } while (0)
    /* ground: optx/rtwo.cxx:154-163 */
    PUSH(ORC_DIFFUSE, mk(.5f, .5f, .5f), 0.f, 0.f, mk(0.f, -1000.f, 0.f), 1000.f, 0);
    for (int a = -half; a < half; a++) {          /* optx/rtwo.cxx:165-207 */
        for (int b = -half; b < half; b++) {
            float select = orc_frand48_next(&rng);
            float cx = (float)a + .9f * orc_frand48_next(&rng);
            float cz = (float)b + .9f * orc_frand48_next(&rng);
            f3 center = mk(cx, .2f, cz);
            if (len(sub(center, mk(4.f, .2f, 0.f))) > .9f) {
                if (select < .8f) {
                    float r1 = orc_frand48_next(&rng), g1 = orc_frand48_next(&rng), b1 = orc_frand48_next(&rng);
                    float r2 = orc_frand48_next(&rng), g2 = orc_frand48_next(&rng), b2 = orc_frand48_next(&rng);
                    PUSH(ORC_DIFFUSE, mk(r1 * r2, g1 * g2, b1 * b2), 0.f, 0.f, center, .2f, 1);
                } else if (select < .95f) {
                    float r1 = rnd_range(&rng, .5f, 1.f), g1 = rnd_range(&rng, .5f, 1.f), b1 = rnd_range(&rng, .5f, 1.f);
                    float fz = rnd_range(&rng, 0.f, .5f);
                    PUSH(ORC_REFLECT, mk(r1, g1, b1), fz, 0.f, center, .2f, 1);
                } else {
                    PUSH(ORC_REFRACT, mk(0.f, 0.f, 0.f), 0.f, 1.5f, center, .2f, 1);
                }
            }
        }
    }
    /* big spheres: optx/rtwo.cxx:209-241 */
    PUSH(ORC_REFRACT, mk(0.f, 0.f, 0.f), 0.f, 1.5f, mk(0.f, 1.f, 0.f), 1.f, 2);
    PUSH(ORC_DIFFUSE, mk(.4f, .2f, .1f), 0.f, 0.f, mk(-4.f, 1.f, 0.f), 1.f, 2);
    PUSH(ORC_REFLECT, mk(.7f, .6f, .5f), 0.f, 0.f, mk(4.f, 1.f, 0.f), 1.f, 2);
#undef PUSH
    return n;
}

void orc_transform_point(const float t[12], const float v[3], float out[3]) {
    /* optx/optics_r.cu:61-65, one rounding per operation, left to right */
    for (int r = 0; r < 3; r++) {
        float acc = v[0] * t[r * 4 + 0];
        acc = acc + v[1] * t[r * 4 + 1];
        acc = acc + v[2] * t[r * 4 + 2];
        out[r] = acc + t[r * 4 + 3];
    }
}

/* ========================================================================= */
/* FP64 closest hit                                                          */
/* ========================================================================= */

/* Moller-Trumbore in double.  Returns 1 and t,u,v on a hit with t in [tmin,tmax]. */
static inline int mt_f64(const float* tri, const double o[3], const double d[3],
                         double tmin, double tmax, double* tt, double* uu, double* vv) {
    double a[3] = {tri[0], tri[1], tri[2]};
    double e1[3] = {(double)tri[3] - a[0], (double)tri[4] - a[1], (double)tri[5] - a[2]};
    double e2[3] = {(double)tri[6] - a[0], (double)tri[7] - a[1], (double)tri[8] - a[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return 0;
    double inv = 1.0 / det;
    double s[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
    double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    if (u < 0.0 || u > 1.0) return 0;
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return 0;
    double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    if (t < tmin || t > tmax) return 0;
    *tt = t; *uu = u; *vv = v;
    return 1;
}

/* keep the closest hit (ties: lower prim) and the second-closest distinct t */
typedef struct { double t, u, v, t2; uint32_t prim; } hit64;

static inline void hit64_offer(hit64* h, double t, double u, double v, uint32_t prim) {
    if (t < h->t || (t == h->t && prim < h->prim)) {
        if (t < h->t) h->t2 = h->t;
        h->t = t; h->u = u; h->v = v; h->prim = prim;
    } else if (t < h->t2 && t != h->t) {
        h->t2 = t;
    } else if (t == h->t && prim != h->prim) {
        /* an exact tie with another triangle: t2 == t marks it undecidable */
        h->t2 = t;
    }
}

typedef struct {
    const float* tris; uint32_t ntris; const orc_bvh* bvh;
    const orc_ray* rays; uint32_t nrays;
    double *t, *u, *v, *t2; uint32_t* prim;
    float *tf, *uf, *vf;
    int tid, nthreads;
} qjob;

static void* brute_worker(void* arg) {
    qjob* j = (qjob*)arg;
    for (uint32_t r = (uint32_t)j->tid; r < j->nrays; r += (uint32_t)j->nthreads) {
        const orc_ray* ray = &j->rays[r];
        double o[3] = {ray->ori[0], ray->ori[1], ray->ori[2]};
        double d[3] = {ray->dir[0], ray->dir[1], ray->dir[2]};
        hit64 h = {INFINITY, 0, 0, INFINITY, 0xFFFFFFFFu};
        for (uint32_t k = 0; k < j->ntris; k++) {
            double tt, uu, vv;
            if (mt_f64(j->tris + 9 * (size_t)k, o, d, ray->tmin, ray->tmax, &tt, &uu, &vv))
                hit64_offer(&h, tt, uu, vv, k);
        }
        j->t[r] = h.t; j->u[r] = h.u; j->v[r] = h.v; j->t2[r] = h.t2; j->prim[r] = h.prim;
    }
    return NULL;
}

static void run_jobs(qjob* proto, int nthreads, void* (*fn)(void*)) {
    if (nthreads <= 0) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    qjob jobs[256];
    for (int i = 0; i < nthreads; i++) {
        jobs[i] = *proto;
        jobs[i].tid = i;
        jobs[i].nthreads = nthreads;
    }
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, fn, &jobs[i]);
    fn(&jobs[0]);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
}

void orc_closest_f64_brute(const float* tris, uint32_t ntris, const orc_ray* rays, uint32_t nrays,
                           double* t, uint32_t* prim, double* u, double* v, double* t2,
                           int nthreads) {
    qjob j;
    memset(&j, 0, sizeof(j));
    j.tris = tris; j.ntris = ntris; j.rays = rays; j.nrays = nrays;
    j.t = t; j.u = u; j.v = v; j.t2 = t2; j.prim = prim;
    run_jobs(&j, nthreads, brute_worker);
}

int orc_decidable(double t, double u, double v, double t2, double tmin, double eps) {
    if (!isfinite(t)) return 1;   /* a miss in FP64 */
    double w = 1.0 - u - v;
    double m = u < v ? u : v;
    if (w < m) m = w;
    if (m < eps) return 0;
    if (!((t2 - t) / t > eps)) return 0;
    if (!((t - tmin) / tmin > eps)) return 0;
    return 1;
}

/* ========================================================================= */
/* CPU BVH (binned SAH, 16 bins, leaves of at most 4 triangles)               */
/* ========================================================================= */

typedef struct { float lo[3], hi[3]; uint32_t left_or_first, count; } bnode;  /* count>0: leaf */

struct orc_bvh {
    const float* tris;
    uint32_t ntris;
    bnode* nodes;
    uint32_t nnodes;
    uint32_t* order;      /* leaf order -> tri index */
};

uint32_t orc_bvh_node_count(const orc_bvh* b) { return b->nnodes; }

typedef struct { float lo[3], hi[3]; } aabb;
static inline void aabb_empty(aabb* b) {
    for (int k = 0; k < 3; k++) { b->lo[k] = FLT_MAX; b->hi[k] = -FLT_MAX; }
}
static inline void aabb_grow_pt(aabb* b, const float* p) {
    for (int k = 0; k < 3; k++) {
        if (p[k] < b->lo[k]) b->lo[k] = p[k];
        if (p[k] > b->hi[k]) b->hi[k] = p[k];
    }
}
static inline void aabb_grow(aabb* b, const aabb* c) {
    for (int k = 0; k < 3; k++) {
        if (c->lo[k] < b->lo[k]) b->lo[k] = c->lo[k];
        if (c->hi[k] > b->hi[k]) b->hi[k] = c->hi[k];
    }
}
static inline float aabb_area(const aabb* b) {
    float dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

orc_bvh* orc_bvh_build(const float* tris, uint32_t ntris) {
    orc_bvh* b = (orc_bvh*)calloc(1, sizeof(orc_bvh));
    b->tris = tris; b->ntris = ntris;
    b->order = (uint32_t*)malloc(sizeof(uint32_t) * (ntris ? ntris : 1));
    b->nodes = (bnode*)malloc(sizeof(bnode) * (2 * (size_t)ntris + 1));
    aabb* pb = (aabb*)malloc(sizeof(aabb) * (ntris ? ntris : 1));
    float* cen = (float*)malloc(sizeof(float) * 3 * (ntris ? ntris : 1));
    for (uint32_t i = 0; i < ntris; i++) {
        aabb_empty(&pb[i]);
        for (int k = 0; k < 3; k++) aabb_grow_pt(&pb[i], tris + 9 * (size_t)i + 3 * k);
        for (int k = 0; k < 3; k++) cen[3 * i + k] = .5f * (pb[i].lo[k] + pb[i].hi[k]);
        b->order[i] = i;
    }
    typedef struct { uint32_t node, first, count; } task;
    task* stack = (task*)malloc(sizeof(task) * 256);
    int sp = 0;
    b->nnodes = 1;
    stack[sp++] = (task){0, 0, ntris};
    enum { NB = 16, MAXLEAF = 4 };
    while (sp > 0) {
        task tk = stack[--sp];
        bnode* nd = &b->nodes[tk.node];
        aabb box, cb;
        aabb_empty(&box); aabb_empty(&cb);
        for (uint32_t i = tk.first; i < tk.first + tk.count; i++) {
            aabb_grow(&box, &pb[b->order[i]]);
            aabb_grow_pt(&cb, cen + 3 * b->order[i]);
        }
        memcpy(nd->lo, box.lo, 12); memcpy(nd->hi, box.hi, 12);
        int best_axis = -1, best_bin = -1;
        float best_cost = FLT_MAX;
        if (tk.count > 1) {
            for (int ax = 0; ax < 3; ax++) {
                float ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0.f)) continue;
                aabb bins[NB]; uint32_t cnt[NB];
                for (int k = 0; k < NB; k++) { aabb_empty(&bins[k]); cnt[k] = 0; }
                float sc = (float)NB / ext;
                for (uint32_t i = tk.first; i < tk.first + tk.count; i++) {
                    uint32_t p = b->order[i];
                    int k = (int)((cen[3 * p + ax] - cb.lo[ax]) * sc);
                    if (k >= NB) k = NB - 1;
                    if (k < 0) k = 0;
                    cnt[k]++; aabb_grow(&bins[k], &pb[p]);
                }
                float rarea[NB]; uint32_t rcnt[NB];
                aabb acc; aabb_empty(&acc); uint32_t c = 0;
                for (int k = NB - 1; k > 0; k--) {
                    aabb_grow(&acc, &bins[k]); c += cnt[k];
                    rarea[k] = aabb_area(&acc); rcnt[k] = c;
                }
                aabb_empty(&acc); c = 0;
                for (int k = 0; k < NB - 1; k++) {
                    aabb_grow(&acc, &bins[k]); c += cnt[k];
                    if (c == 0 || rcnt[k + 1] == 0) continue;
                    float cost = aabb_area(&acc) * (float)c + rarea[k + 1] * (float)rcnt[k + 1];
                    if (cost < best_cost) { best_cost = cost; best_axis = ax; best_bin = k; }
                }
            }
        }
        float leaf_cost = aabb_area(&box) * (float)tk.count;
        int make_leaf = tk.count <= 1 ||
                        (tk.count <= MAXLEAF && (best_axis < 0 || leaf_cost <= best_cost + aabb_area(&box)));
        if (make_leaf && tk.count <= MAXLEAF) {
            nd->left_or_first = tk.first;
            nd->count = tk.count;
            continue;
        }
        uint32_t mid;
        if (best_axis >= 0) {
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            float sc = (float)NB / ext;
            uint32_t i = tk.first, j = tk.first + tk.count;
            while (i < j) {
                uint32_t p = b->order[i];
                int k = (int)((cen[3 * p + best_axis] - cb.lo[best_axis]) * sc);
                if (k >= NB) k = NB - 1;
                if (k < 0) k = 0;
                if (k <= best_bin) i++;
                else { j--; uint32_t tmp = b->order[i]; b->order[i] = b->order[j]; b->order[j] = tmp; }
            }
            mid = i;
        } else {
            mid = tk.first + tk.count / 2;   /* coincident centroids: split by index */
        }
        if (mid == tk.first || mid == tk.first + tk.count) mid = tk.first + tk.count / 2;
        uint32_t l = b->nnodes, r = b->nnodes + 1;
        b->nnodes += 2;
        nd = &b->nodes[tk.node];
        nd->left_or_first = l;
        nd->count = 0;
        if (sp + 2 > 256) { /* cannot happen for sane inputs; degrade to a fat leaf */
            nd->left_or_first = tk.first; nd->count = tk.count; b->nnodes -= 2; continue;
        }
        stack[sp++] = (task){r, mid, tk.first + tk.count - mid};
        stack[sp++] = (task){l, tk.first, mid - tk.first};
    }
    free(stack); free(pb); free(cen);
    return b;
}

void orc_bvh_free(orc_bvh* b) {
    if (!b) return;
    free(b->nodes); free(b->order); free(b);
}

/* slab test in double, padded so that it never rejects a box the exact ray meets */
static inline int box_f64(const bnode* n, const double o[3], const double inv[3],
                          double tmin, double tmax) {
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        double a = ((double)n->lo[k] - o[k]) * inv[k];
        double b = ((double)n->hi[k] - o[k]) * inv[k];
        if (a != a) a = -INFINITY;   /* 0*inf: ray on the slab plane */
        if (b != b) b = INFINITY;
        double lo = a < b ? a : b, hi = a < b ? b : a;
        if (lo > t0) t0 = lo;
        if (hi < t1) t1 = hi;
    }
    return t0 <= t1 * (1.0 + 1e-9) + 1e-300;
}

static void* bvh64_worker(void* arg) {
    qjob* j = (qjob*)arg;
    const orc_bvh* b = j->bvh;
    uint32_t stack[128];
    for (uint32_t r = (uint32_t)j->tid; r < j->nrays; r += (uint32_t)j->nthreads) {
        const orc_ray* ray = &j->rays[r];
        double o[3] = {ray->ori[0], ray->ori[1], ray->ori[2]};
        double d[3] = {ray->dir[0], ray->dir[1], ray->dir[2]};
        double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
        hit64 h = {INFINITY, 0, 0, INFINITY, 0xFFFFFFFFu};
        int sp = 0;
        if (b->ntris) stack[sp++] = 0;
        while (sp > 0) {
            const bnode* n = &b->nodes[stack[--sp]];
            /* prune against the second-closest hit so that t2 stays exact */
            if (!box_f64(n, o, inv, ray->tmin, h.t2 < ray->tmax ? h.t2 : ray->tmax)) continue;
            if (n->count) {
                for (uint32_t i = n->left_or_first; i < n->left_or_first + n->count; i++) {
                    uint32_t p = b->order[i];
                    double tt, uu, vv;
                    if (mt_f64(b->tris + 9 * (size_t)p, o, d, ray->tmin, ray->tmax, &tt, &uu, &vv))
                        hit64_offer(&h, tt, uu, vv, p);
                }
            } else {
                stack[sp++] = n->left_or_first + 1;
                stack[sp++] = n->left_or_first;
            }
        }
        j->t[r] = h.t; j->u[r] = h.u; j->v[r] = h.v; j->t2[r] = h.t2; j->prim[r] = h.prim;
    }
    return NULL;
}

void orc_closest_f64_bvh(const orc_bvh* b, const orc_ray* rays, uint32_t nrays,
                         double* t, uint32_t* prim, double* u, double* v, double* t2,
                         int nthreads) {
    qjob j;
    memset(&j, 0, sizeof(j));
    j.bvh = b; j.rays = rays; j.nrays = nrays;
    j.t = t; j.u = u; j.v = v; j.t2 = t2; j.prim = prim;
    run_jobs(&j, nthreads, bvh64_worker);
}

/* ========================================================================= */
/* FP32 watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1), 2013)   */
/* ========================================================================= */

typedef struct {
    float o[3], d[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float inv[3];
    float tmin;
} wray;

static inline void wray_init(wray* r, const float o[3], const float d[3], float tmin) {
    memcpy(r->o, o, 12); memcpy(r->d, d, 12);
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    /* Woop et al. swap kx and ky when d[kz] < 0 to keep the winding; that
     * negates U, V, W, det and T alike and leaves t, u, v bit for bit the
     * same, and no test here depends on the winding, so no swap */
    r->kx = kx; r->ky = ky; r->kz = kz;
    r->Sz = 1.f / d[kz];
    r->Sx = d[kx] * r->Sz;     /* shear by the reciprocal (one division per ray) */
    r->Sy = d[ky] * r->Sz;
    for (int k = 0; k < 3; k++) r->inv[k] = 1.f / (d[k] != 0.f ? d[k] : 1e-30f);
    r->tmin = tmin;
}

/* returns 1 with t,u,v if the triangle is hit with tmin < t; u,v weight v1,v2 */
static inline int watertight(const wray* r, const float* tri, float* tt, float* uu, float* vv) {
    float A[3], B[3], C[3];
    for (int k = 0; k < 3; k++) {
        A[k] = tri[k] - r->o[k];
        B[k] = tri[3 + k] - r->o[k];
        C[k] = tri[6 + k] - r->o[k];
    }
    const int kx = r->kx, ky = r->ky, kz = r->kz;
    float Ax = A[kx] - r->Sx * A[kz], Ay = A[ky] - r->Sy * A[kz];
    float Bx = B[kx] - r->Sx * B[kz], By = B[ky] - r->Sy * B[kz];
    float Cx = C[kx] - r->Sx * C[kz], Cy = C[ky] - r->Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return 0;
    float det = U + V + W;
    if (det == 0.f) return 0;
    float Az = r->Sz * A[kz], Bz = r->Sz * B[kz], Cz = r->Sz * C[kz];
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.f / det;
    float t = T * rcp;
    if (!(t > r->tmin)) return 0;
    *tt = t; *uu = V * rcp; *vv = W * rcp;
    return 1;
}

static inline int box_f32(const bnode* n, const wray* r, float tmax) {
    float t0 = r->tmin, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        float a = (n->lo[k] - r->o[k]) * r->inv[k];
        float b = (n->hi[k] - r->o[k]) * r->inv[k];
        float lo = a < b ? a : b, hi = a < b ? b : a;
        if (lo > t0) t0 = lo;
        if (hi < t1) t1 = hi;
    }
    /* conservative: pad the far distance by 2*gamma(3) (Ize 2013) */
    return t0 <= t1 * 1.0000004f;
}

typedef struct { float t, u, v; uint32_t prim; } hit32;

static void trace_f32(const orc_bvh* b, const wray* r, float tmax, hit32* h) {
    uint32_t stack[128];
    int sp = 0;
    h->t = tmax; h->u = h->v = 0.f; h->prim = 0xFFFFFFFFu;
    if (b->ntris) stack[sp++] = 0;
    while (sp > 0) {
        const bnode* n = &b->nodes[stack[--sp]];
        if (!box_f32(n, r, h->t)) continue;
        if (n->count) {
            for (uint32_t i = n->left_or_first; i < n->left_or_first + n->count; i++) {
                uint32_t p = b->order[i];
                float tt, uu, vv;
                if (watertight(r, b->tris + 9 * (size_t)p, &tt, &uu, &vv)) {
                    if (tt < h->t || (tt == h->t && p < h->prim)) {
                        h->t = tt; h->u = uu; h->v = vv; h->prim = p;
                    }
                }
            }
        } else {
            stack[sp++] = n->left_or_first + 1;
            stack[sp++] = n->left_or_first;
        }
    }
}

static void* bvh32_worker(void* arg) {
    qjob* j = (qjob*)arg;
    for (uint32_t i = (uint32_t)j->tid; i < j->nrays; i += (uint32_t)j->nthreads) {
        const orc_ray* ray = &j->rays[i];
        wray r;
        wray_init(&r, ray->ori, ray->dir, ray->tmin);
        hit32 h;
        trace_f32(j->bvh, &r, ray->tmax, &h);
        if (h.prim == 0xFFFFFFFFu) h.t = INFINITY;
        j->tf[i] = h.t; j->uf[i] = h.u; j->vf[i] = h.v; j->prim[i] = h.prim;
    }
    return NULL;
}

void orc_closest_f32_bvh(const orc_bvh* b, const orc_ray* rays, uint32_t nrays,
                         float* t, uint32_t* prim, float* u, float* v, int nthreads) {
    qjob j;
    memset(&j, 0, sizeof(j));
    j.bvh = b; j.rays = rays; j.nrays = nrays;
    j.tf = t; j.uf = u; j.vf = v; j.prim = prim;
    run_jobs(&j, nthreads, bvh32_worker);
}

/* ========================================================================= */
/* Render with the recursive GPU program semantics                           */
/*   raygen  optx/camera_r.cu:23-121, miss optx/camera_r.cu:123-147,          */
/*   closest hits optx/optics_r.cu:22-431                                     */
/* ========================================================================= */

typedef struct { f3 normal, albedo; } dgv_t;   /* optx/rtwo.h:90-93 */

/* optx/optics_r.cu:235-236 and :392: reflect a unit direction about N */
static inline f3 reflect_unit(f3 d1, f3 N) { return sub(d1, scl(2.f * dot(d1, N), N)); }
/* optx/optics_r.cu:394-397: refract a unit direction, ior ratio */
static inline f3 refract_unit(f3 d1, f3 N, float ratio) {
    float theta = fminf(dot(neg(d1), N), 1.f);
    f3 perpen = scl(ratio, add(d1, scl(theta, N)));
    f3 parall = scl(-sqrtf(fabsf(1.f - dot(perpen, perpen))), N);
    return add(perpen, parall);
}
/* optx/optics_r.cu:387-388; the fifth power as x^2 x^2 x (three roundings)
 * where the reference calls powf */
static inline float schlick_f(float cos_t, float ratio) {
    float r0 = (1.f - ratio) / (1.f + ratio); r0 = r0 * r0;
    const float x = 1.f - cos_t, x2 = x * x;
    return r0 + (1.f - r0) * ((x2 * x2) * x);
}
/* optx/camera_r.cu:135-140 */
static inline f3 sky_f(f3 dir, f3 amb) {
    f3 unit = unitv(dir);
    float t = .5f * (unit.y + 1.f);
    return add(scl(1.f - t, mk(1.f, 1.f, 1.f)), scl(t, amb));
}

void orc_reflect(const float d[3], const float n[3], float out[3]) { st3(out, reflect_unit(unitv(ld3(d)), ld3(n))); }
void orc_refract(const float d[3], const float n[3], float ratio, float out[3]) {
    st3(out, refract_unit(unitv(ld3(d)), ld3(n), ratio));
}
float orc_schlick(float cos_t, float ratio) { return schlick_f(cos_t, ratio); }
void orc_sky(const float d[3], const float amb[3], float out[3]) { st3(out, sky_f(ld3(d), ld3(amb))); }

typedef struct {
    const orc_scene* s;
    uint32_t depth;
    orc_frand48* rng;
    dgv_t* dgv;
} tctx;

/* One sample path: the recursion of the closest-hit programs
 * (optx/optics_r.cu:22-431 with the miss program optx/camera_r.cu:123-147)
 * unrolled into a loop.  The reference multiplies the albedos on the way back
 * up the recursion (col = alb * child); the loop multiplies them into a
 * throughput on the way down, the order the GPU kernel uses (same factors,
 * other association), so that both round alike.  Guide values: the recursive
 * programs set them after the recursion returns, where still null
 * (optx/optics_r.cu:143-146, :280-283, util::isnull optx/util_gpu.h:39), so
 * the deepest diffuse/metal hit's normal and the deepest non-zero albedo win.
 * With s->iterative the loop follows optx/camera_i.cu:75-95 and
 * optx/optics_i.cu instead: the hit of the depth-th ray is shaded like any
 * other and the sample keeps the throughput as its colour, and the guides are
 * set on the way down where still null (optx/optics_i.cu:98-101): the first
 * hit's normal and the first non-zero albedo.
 * Returns the colour; *depth = rays traced (payload 5). */
static f3 trace_path(const tctx* c, f3 ori, f3 dir, uint32_t* depth) {
    const orc_scene* s = c->s;
    const int iter = s->iterative != 0;
    f3 thr = mk(1.f, 1.f, 1.f);
    uint32_t dep = 0;
    for (;;) {
        wray r;
        float o[3] = {ori.x, ori.y, ori.z}, d[3] = {dir.x, dir.y, dir.z};
        wray_init(&r, o, d, 1e-3f);
        hit32 h;
        trace_f32(s->bvh, &r, 1e16f, &h);
        dep++;                                             /* optx/optics_r.cu:31, camera_r.cu:146 */
        if (h.prim == 0xFFFFFFFFu) {
            /* __miss__ambient, optx/camera_r.cu:131-146 */
            *depth = dep;
            return mul(thr, sky_f(dir, ld3(s->ambient)));
        }
        if (!iter && c->depth == dep) { *depth = dep; return mk(0.f, 0.f, 0.f); }   /* :34-42 */
        const orc_thing* th = &s->things[s->tri_inst[h.prim]];
        const float* tri = s->tris + 9 * (size_t)h.prim;
        f3 A = ld3(tri), B = ld3(tri + 3), C = ld3(tri + 6);
        float u = h.u, v = h.v, w = 1.f - u - v;          /* :78-81 */
        f3 hit = add(add(scl(w, A), scl(u, B)), scl(v, C)); /* :84 */
        f3 N = unitv(cross(sub(B, A), sub(C, A)));         /* :88 */
        if (dot(dir, N) > 0.f) N = neg(N);                 /* :89-90 */
        f3 nd;
        if (th->type != ORC_REFRACT) {
            f3 alb = ld3(th->albedo);
            if (iter) {                                    /* optx/optics_i.cu:98-101, :173-176 */
                if (isnull3(c->dgv->normal)) c->dgv->normal = N;
                if (isnull3(c->dgv->albedo)) c->dgv->albedo = alb;
            } else {                                       /* optx/optics_r.cu:143-146, :280-283 */
                c->dgv->normal = N;
                if (!isnull3(alb)) c->dgv->albedo = alb;
            }
        }
        if (th->type == ORC_DIFFUSE) {
            nd = add(N, rnd_on_sphere(c->rng));            /* :108 */
            thr = mul(thr, ld3(th->albedo));               /* :135-138 */
        } else if (th->type == ORC_REFLECT) {
            f3 d1 = unitv(dir);                            /* :235 */
            f3 refl = reflect_unit(d1, N);                 /* :236 */
            nd = add(refl, scl(th->fuzz, rnd_in_sphere(c->rng)));  /* :238 */
            if (!(dot(nd, N) > 0)) { *depth = dep; return mk(0.f, 0.f, 0.f); }   /* :242 */
            thr = mul(thr, ld3(th->albedo));
        } else {
            f3 d1 = unitv(dir);                            /* :372 */
            float cos_t = fminf(dot(neg(d1), N), 1.f);     /* :373 */
            float sin_t = sqrtf(1.f - cos_t * cos_t);      /* :374 */
            f3 O = sub(hit, ld3(th->center));              /* :376-381 */
            float ratio = 0.f > dot(dir, O) ? 1.f / th->index : th->index;   /* :382-384 */
            int cannot = ratio * sin_t > 1.f;              /* :385 */
            float schlick = schlick_f(cos_t, ratio);       /* :387-388 */
            if (cannot || schlick > orc_frand48_next(c->rng))   /* :391 */
                nd = reflect_unit(d1, N);
            else
                nd = refract_unit(d1, N, ratio);           /* :394-397 */
            /* :425-428 multiply by (1,1,1): exact, left out */
        }
        if (iter && c->depth == dep) { *depth = dep; return thr; }   /* optx/camera_i.cu:75-95 */
        ori = hit;
        dir = nd;
    }
}

typedef struct {
    const orc_scene* s; const orc_camera* cam;
    uint32_t w, h, spp, depth, y0, y1, seed_base;
    float *raw, *normals, *albedos, *sumsq;
    uint32_t* rpp;
    uint64_t rays;
    int tid, nthreads;
} rjob;

static inline float clampf(float x, float lo, float hi) { return lo > x ? lo : x > hi ? hi : x; }

static void* render_worker(void* arg) {
    rjob* j = (rjob*)arg;
    uint64_t rays = 0;
    for (uint32_t y = j->y0 + (uint32_t)j->tid; y < j->y1; y += (uint32_t)j->nthreads) {
        for (uint32_t x = 0; x < j->w; x++) {
            const uint32_t pix = j->w * y + x;          /* optx/camera_r.cu:35 */
            orc_frand48 rng;
            orc_frand48_init(&rng, pix + j->seed_base); /* optx/camera_r.cu:42-43 */
            f3 color = mk(0, 0, 0), normal = mk(0, 0, 0), albedo = mk(0, 0, 0), sq = mk(0, 0, 0);
            uint32_t rpp = 0;
            for (uint32_t i = 0; i < j->spp; i++) {     /* optx/camera_r.cu:66-109 */
                float s, t;
                pixel_st(x, y, j->w, j->h, &rng, &s, &t);
                float o[3], d[3];
                orc_camera_ray(j->cam, s, t, &rng, o, d);
                dgv_t dgv = {mk(0, 0, 0), mk(0, 0, 0)};
                tctx c = {j->s, j->depth, &rng, &dgv};
                uint32_t depth = 0;
                f3 col = trace_path(&c, ld3(o), ld3(d), &depth);
                color = add(color, col);
                sq = add(sq, mul(col, col));
                rpp = rpp + depth;
                normal = add(normal, dgv.normal);
                albedo = add(albedo, dgv.albedo);
            }
            float inv = 1.f / (float)j->spp;            /* optx/v.h:44: 1/b * a */
            f3 mean = scl(inv, color);
            float* px = j->raw + 3 * (size_t)pix;       /* optx/camera_r.cu:112 */
            px[0] = clampf(mean.x, 0.f, 1.f); px[1] = clampf(mean.y, 0.f, 1.f); px[2] = clampf(mean.z, 0.f, 1.f);
            if (j->rpp) j->rpp[pix] = rpp;              /* optx/camera_r.cu:114 */
            if (j->normals) st3(j->normals + 3 * (size_t)pix, scl(inv, normal));   /* :117-120 */
            if (j->albedos) st3(j->albedos + 3 * (size_t)pix, scl(inv, albedo));
            if (j->sumsq) st3(j->sumsq + 3 * (size_t)pix, sq);
            rays += rpp;
        }
    }
    j->rays = rays;
    return NULL;
}

uint64_t orc_render(const orc_scene* s, const orc_camera* cam, uint32_t w, uint32_t h,
                    uint32_t spp, uint32_t depth, uint32_t y0, uint32_t y1, int nthreads,
                    float* raw, uint32_t* rpp, float* normals, float* albedos, float* sumsq,
                    uint32_t seed_base) {
    if (nthreads <= 0) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (y1 > h) y1 = h;
    pthread_t th[256];
    rjob jobs[256];
    for (int i = 0; i < nthreads; i++) {
        rjob jb = {s, cam, w, h, spp, depth, y0, y1, seed_base, raw, normals, albedos, sumsq, rpp, 0, i, nthreads};
        jobs[i] = jb;
    }
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, render_worker, &jobs[i]);
    render_worker(&jobs[0]);
    uint64_t total = jobs[0].rays;
    for (int i = 1; i < nthreads; i++) { pthread_join(th[i], NULL); total += jobs[i].rays; }
    return total;
}

/* ========================================================================= */
/* Frozen-structure evaluation (finite-difference checker for material      */
/* gradients).  A sample path is traced exactly as trace_path does it (same  */
/* random draws, same decisions) with the scene's own materials, its        */
/* vertices are recorded, and its radiance is then evaluated in FP64 along   */
/* the same vertices with other material parameters: directions follow the  */
/* recorded normals, random vectors and choices; nothing is re-traced.       */
/* Scatter formulas: optx/optics_r.cu:108 (diffuse), :235-242 (metal),       */
/* :372-398 (dielectric); sky optx/camera_r.cu:135-140.                       */
/* ========================================================================= */

typedef struct { int type; uint32_t thing; f3 N; f3 rv; float ratio; int refl; } fvert;

/* returns 1 if the path reaches the sky, 0 if it ends black */
static int frozen_trace(const orc_scene* s, uint32_t maxd, orc_frand48* rng, f3 ori, f3 dir, fvert* vs, int* nv) {
    uint32_t dep = 0;
    *nv = 0;
    for (;;) {
        wray r;
        float o[3] = {ori.x, ori.y, ori.z}, d[3] = {dir.x, dir.y, dir.z};
        wray_init(&r, o, d, 1e-3f);
        hit32 h;
        trace_f32(s->bvh, &r, 1e16f, &h);
        if (h.prim == 0xFFFFFFFFu) return 1;
        dep++;
        if (dep == maxd) return 0;
        const uint32_t ti = s->tri_inst[h.prim];
        const orc_thing* th = &s->things[ti];
        const float* tri = s->tris + 9 * (size_t)h.prim;
        f3 A = ld3(tri), B = ld3(tri + 3), C = ld3(tri + 6);
        float u = h.u, v = h.v, w = 1.f - u - v;
        f3 hit = add(add(scl(w, A), scl(u, B)), scl(v, C));
        f3 N = unitv(cross(sub(B, A), sub(C, A)));
        if (dot(dir, N) > 0.f) N = neg(N);
        fvert* fv = &vs[(*nv)++];
        fv->type = th->type; fv->thing = ti; fv->N = N; fv->ratio = 0.f; fv->refl = 0;
        f3 nd;
        if (th->type == ORC_DIFFUSE) {
            fv->rv = rnd_on_sphere(rng);
            nd = add(N, fv->rv);
        } else if (th->type == ORC_REFLECT) {
            f3 d1 = unitv(dir);
            fv->rv = rnd_in_sphere(rng);
            nd = add(reflect_unit(d1, N), scl(th->fuzz, fv->rv));
            if (!(dot(nd, N) > 0)) return 0;
        } else {
            f3 d1 = unitv(dir);
            float cos_t = fminf(dot(neg(d1), N), 1.f);
            float sin_t = sqrtf(1.f - cos_t * cos_t);
            f3 O = sub(hit, ld3(th->center));
            float ratio = 0.f > dot(dir, O) ? 1.f / th->index : th->index;
            int cannot = ratio * sin_t > 1.f;
            float schlick = schlick_f(cos_t, ratio);
            fv->ratio = ratio;
            if (cannot || schlick > orc_frand48_next(rng)) { fv->refl = 1; nd = reflect_unit(d1, N); }
            else nd = refract_unit(d1, N, ratio);
        }
        ori = hit;
        dir = nd;
    }
}

static void frozen_eval(const fvert* vs, int nv, f3 dir0, const float* params, const float amb[3], double out[3]) {
    double d[3] = {dir0.x, dir0.y, dir0.z}, thr[3] = {1, 1, 1};
    for (int k = 0; k < nv; k++) {
        const fvert* v = &vs[k];
        const float* p = params + 4 * (size_t)v->thing;
        const double N[3] = {v->N.x, v->N.y, v->N.z};
        if (v->type == ORC_DIFFUSE) {
            for (int c = 0; c < 3; c++) thr[c] *= p[c];
            d[0] = N[0] + v->rv.x; d[1] = N[1] + v->rv.y; d[2] = N[2] + v->rv.z;
            continue;
        }
        const double il = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double u[3] = {d[0] * il, d[1] * il, d[2] * il};
        const double un = u[0] * N[0] + u[1] * N[1] + u[2] * N[2];
        if (v->type == ORC_REFLECT) {
            const double rv[3] = {v->rv.x, v->rv.y, v->rv.z};
            for (int c = 0; c < 3; c++) { d[c] = u[c] - 2 * un * N[c] + (double)p[3] * rv[c]; thr[c] *= p[c]; }
        } else if (v->refl) {
            for (int c = 0; c < 3; c++) d[c] = u[c] - 2 * un * N[c];
        } else {
            const double th = fmin(-un, 1.0), ratio = v->ratio;
            double pp[3], q = 1;
            for (int c = 0; c < 3; c++) { pp[c] = ratio * (u[c] + th * N[c]); q -= pp[c] * pp[c]; }
            const double sq = sqrt(fabs(q));
            for (int c = 0; c < 3; c++) d[c] = pp[c] - sq * N[c];
        }
    }
    const double il = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double t = .5 * (d[1] * il + 1.0);
    for (int c = 0; c < 3; c++) out[c] = thr[c] * ((1.0 - t) + t * amb[c]);
}

typedef struct {
    const orc_scene* s; const orc_camera* cam;
    uint32_t w, h, spp, depth, y0, y1;
    const float* params;
    double* raw;
    int tid, nthreads;
} fjob;

static void* frozen_worker(void* arg) {
    fjob* j = (fjob*)arg;
    fvert vs[64];
    for (uint32_t y = j->y0 + (uint32_t)j->tid; y < j->y1; y += (uint32_t)j->nthreads) {
        for (uint32_t x = 0; x < j->w; x++) {
            const uint32_t pix = j->w * y + x;
            orc_frand48 rng;
            orc_frand48_init(&rng, pix);
            double acc[3] = {0, 0, 0};
            for (uint32_t i = 0; i < j->spp; i++) {
                float s, t;
                pixel_st(x, y, j->w, j->h, &rng, &s, &t);
                float o[3], d[3];
                orc_camera_ray(j->cam, s, t, &rng, o, d);
                int nv = 0;
                if (frozen_trace(j->s, j->depth, &rng, ld3(o), ld3(d), vs, &nv)) {
                    double col[3];
                    frozen_eval(vs, nv, ld3(d), j->params, j->s->ambient, col);
                    for (int c = 0; c < 3; c++) acc[c] += col[c];
                }
            }
            for (int c = 0; c < 3; c++) {
                const double m = acc[c] / j->spp;
                j->raw[3 * (size_t)pix + c] = m < 0 ? 0 : m > 1 ? 1 : m;
            }
        }
    }
    return NULL;
}

void orc_render_frozen(const orc_scene* s, const orc_camera* cam, uint32_t w, uint32_t h, uint32_t spp,
                       uint32_t depth, uint32_t y0, uint32_t y1, int nthreads, const float* params, double* raw) {
    if (nthreads <= 0) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (y1 > h) y1 = h;
    if (depth > 64) depth = 64;
    pthread_t th[256];
    fjob jobs[256];
    for (int i = 0; i < nthreads; i++) {
        fjob jb = {s, cam, w, h, spp, depth, y0, y1, params, raw, i, nthreads};
        jobs[i] = jb;
    }
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, frozen_worker, &jobs[i]);
    frozen_worker(&jobs[0]);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
}

/* ========================================================================= */
/* Restatement of rtow.cxx (FP64, analytic spheres, glibc rand()).            */
/* Operation order follows the reference expressions one for one:            */
/* v.h, ray.h, util.h, sphere.h, things.h, camera.h, optics.h, rtow.cxx.       */
/* ========================================================================= */

typedef struct { double x, y, z; } d3;
static inline d3 dk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 dadd(d3 a, d3 b) { return dk(a.x + b.x, a.y + b.y, a.z + b.z); }      /* v.h:42 */
static inline d3 dsub(d3 a, d3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }      /* v.h:43 */
static inline d3 dmul(d3 a, d3 b) { return dk(a.x * b.x, a.y * b.y, a.z * b.z); }      /* v.h:44 */
static inline d3 dscl(double t, d3 v) { return dk(t * v.x, t * v.y, t * v.z); }        /* v.h:45 */
static inline d3 ddiv(d3 v, double t) { return dscl(1 / t, v); }                       /* v.h:46 */
static inline d3 dneg(d3 a) { return dk(-a.x, -a.y, -a.z); }                           /* v.h:18 */
static inline double ddot(d3 u, d3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }    /* v.h:48 */
static inline double dsq(d3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }           /* v.h:28 */
static inline double dlen(d3 v) { return sqrt(dsq(v)); }                               /* v.h:27 */
static inline d3 dcross(d3 u, d3 v) {                                                  /* v.h:49 */
    return dk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline d3 dunit(d3 v) { return ddiv(v, dlen(v)); }                               /* v.h:50 */

static int g_r2l = 1;
/* util.h:12-13 */
static inline double rrnd(void) { return rand() / (RAND_MAX + 1.); }
static inline double rrnd2(double lo, double hi) { return lo + rrnd() * (hi - lo); }
/* a V built from three calls: V( f(), f(), f() ) with the compiler's order */
static d3 rv3(double lo, double hi, int ranged) {
    double a, b, c;
    if (g_r2l) { c = ranged ? rrnd2(lo, hi) : rrnd(); b = ranged ? rrnd2(lo, hi) : rrnd(); a = ranged ? rrnd2(lo, hi) : rrnd(); }
    else { a = ranged ? rrnd2(lo, hi) : rrnd(); b = ranged ? rrnd2(lo, hi) : rrnd(); c = ranged ? rrnd2(lo, hi) : rrnd(); }
    return dk(a, b, c);
}
typedef struct { d3 c; double r; int type; d3 albedo; double fuzz, index; } dsphere;

typedef struct { dsphere* s; int n, cap; } dthings;
static void dth_add(dthings* t, d3 c, double r, int type, d3 alb, double fuzz, double index) {
    if (t->n == t->cap) { t->cap = t->cap ? 2 * t->cap : 64; t->s = (dsphere*)realloc(t->s, sizeof(dsphere) * t->cap); }
    dsphere sp = {c, r, type, alb, fuzz, index};
    t->s[t->n++] = sp;
}

/* rtow.cxx:51-80 */
static dthings rtow_scene(void) {
    dthings s = {0, 0, 0};
    dth_add(&s, dk(0, -1000, 0), 1000., ORC_DIFFUSE, dk(.5, .5, .5), 0, 0);
    for (int a = -11; a < 11; a++) {
        for (int b = -11; b < 11; b++) {
            double select = rrnd();
            double cx, cz;   /* P center( a+.9*rnd(), .2, b+.9*rnd() ) */
            if (g_r2l) { cz = b + .9 * rrnd(); cx = a + .9 * rrnd(); }
            else { cx = a + .9 * rrnd(); cz = b + .9 * rrnd(); }
            d3 center = dk(cx, .2, cz);
            if (dlen(dsub(center, dk(4, .2, 0))) > .9) {
                if (select < .8) {
                    d3 p, q;   /* C::rnd()*C::rnd() */
                    if (g_r2l) { q = rv3(0, 0, 0); p = rv3(0, 0, 0); } else { p = rv3(0, 0, 0); q = rv3(0, 0, 0); }
                    dth_add(&s, center, .2, ORC_DIFFUSE, dmul(p, q), 0, 0);
                } else if (select < .95) {
                    d3 alb = rv3(.5, 1., 1);
                    double fuzz = rrnd2(0, .5);
                    dth_add(&s, center, .2, ORC_REFLECT, alb, fuzz, 0);
                } else {
                    dth_add(&s, center, .2, ORC_REFRACT, dk(0, 0, 0), 0, 1.5);
                }
            }
        }
    }
    dth_add(&s, dk(0, 1, 0), 1., ORC_REFRACT, dk(0, 0, 0), 0, 1.5);
    dth_add(&s, dk(-4, 1, 0), 1., ORC_DIFFUSE, dk(.4, .2, .1), 0, 0);
    dth_add(&s, dk(4, 1, 0), 1., ORC_REFLECT, dk(.7, .6, .5), 0, 0);
    return s;
}

uint32_t orc_rtow_scene(int argorder, orc_thing* out, uint32_t cap) {
    g_r2l = argorder;
    srand(1);
    dthings s = rtow_scene();
    uint32_t n = (uint32_t)s.n;
    for (uint32_t i = 0; i < n && i < cap; i++) {
        dsphere* p = &s.s[i];
        set_thing(&out[i], p->type, mk((float)p->albedo.x, (float)p->albedo.y, (float)p->albedo.z),
                  (float)p->fuzz, (float)p->index, mk((float)p->c.x, (float)p->c.y, (float)p->c.z),
                  (float)p->r, i == 0 ? 0 : (p->r == 1. ? 2 : 1));
    }
    free(s.s);
    return n;
}

/* Random draws of the tracer (util.h:12): glibc rand() as the reference
 * draws them, or a pixel's own drand48 stream (a = 0x5DEECE66D, c = 0xB mod
 * 2^48: the value erand48 returns, never 1) for the multi-threaded runs */
typedef struct { int r2l; int glibc; uint64_t st; uint64_t rays; } rtrng;
static inline double rt_glibc_draw_d(void) { return rand() / (RAND_MAX + 1.); }
/* FP32: the top 24 of rand()'s 31 bits, so that the float stays below 1 */
static inline float rt_glibc_draw_f(void) { return (float)(rand() >> 7) * 0x1p-24f; }
static inline void rt_stream_next(uint64_t* st) { *st = (0x5DEECE66Dull * *st + 0xBull) & 0xFFFFFFFFFFFFull; }
static inline double rt_stream_draw_d(uint64_t* st) { rt_stream_next(st); return (double)*st * 0x1p-48; }
static inline float rt_stream_draw_f(uint64_t* st) { rt_stream_next(st); return (float)(*st >> 24) * 0x1p-24f; }
static inline void rt_stream_seed(uint64_t* st, uint32_t seed) {
    *st = seed;
    for (int i = 0; i < 16; i++) rt_stream_next(st);
}

#define RT_R double
#define RT_S _d
#define RT_SQRT sqrt
#define RT_FABS fabs
#define RT_FMIN fmin
#define RT_POW pow
#define RT_TAN tan
#include "rtow.inc"
#undef RT_R
#undef RT_S
#undef RT_SQRT
#undef RT_FABS
#undef RT_FMIN
#undef RT_POW
#undef RT_TAN

#define RT_R float
#define RT_S _f
#define RT_SQRT sqrtf
#define RT_FABS fabsf
#define RT_FMIN fminf
#define RT_POW(x, n) powf((x), (float)(n))
#define RT_TAN tanf
#include "rtow.inc"
#undef RT_R
#undef RT_S
#undef RT_SQRT
#undef RT_FABS
#undef RT_FMIN
#undef RT_POW
#undef RT_TAN

uint64_t orc_rtow_render(int w, int h, int spp, int depth, int argorder, uint8_t* rgb) {
    return rtow_run_d(w, h, spp, depth, argorder, 0, 0, (uint64_t)w * (uint64_t)h, rgb, NULL, NULL, NULL);
}

uint64_t orc_rtow_rays(int w, int h, int spp, int depth, int argorder, uint64_t npix) {
    return rtow_run_d(w, h, spp, depth, argorder, 0, 0, npix, NULL, NULL, NULL, NULL);
}

uint64_t orc_rtow_render_ex(const orc_rtow_opts* o, int w, int h, int spp, int depth, uint8_t* rgb, double* mean,
                            double* sumsq, uint32_t* rpp) {
    const uint64_t all = (uint64_t)w * (uint64_t)h;
    if (o->fp32) return rtow_run_f(w, h, spp, depth, o->argorder, o->nthreads, o->seed_base, all, rgb, mean, sumsq, rpp);
    return rtow_run_d(w, h, spp, depth, o->argorder, o->nthreads, o->seed_base, all, rgb, mean, sumsq, rpp);
}

/* ========================================================================= */
/* Post-process: optx/postproc.cu:1-61                                        */
/* ========================================================================= */

static inline float srgb1(float x) {   /* optx/postproc.cu:3 */
    return x < .0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - .055f;
}

void orc_pp_srgb(const float* src, uint8_t* dst, int w, int h) {
    for (int p = 0; p < w * h; p++) {
        for (int k = 0; k < 3; k++) dst[4 * p + k] = (uint8_t)(srgb1(src[3 * p + k]) * 255);
        dst[4 * p + 3] = 255;
    }
}

void orc_pp_srgb_f64(const float* src, double* dst, int w, int h) {
    for (int p = 0; p < w * h; p++)
        for (int k = 0; k < 3; k++) {
            double x = src[3 * p + k];
            dst[3 * p + k] = 255.0 * (x < .0031308 ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - .055);
        }
}

void orc_pp_none(const float* src, uint8_t* dst, int w, int h) {
    for (int p = 0; p < w * h; p++) {
        for (int k = 0; k < 3; k++) dst[4 * p + k] = (uint8_t)(src[3 * p + k] * 255);
        dst[4 * p + 3] = 255;
    }
}
