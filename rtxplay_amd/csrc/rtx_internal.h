// rtx_internal.h -- device data layouts and helpers shared by the kernels.
//
// Layout in HBM of a built scene (what a traversable handle points at):
//   nodes : 4-wide BVH nodes, 128 B each = 8 x float4, structure of arrays
//           over the 4 children: lo.x[4], hi.x[4], lo.y[4], hi.y[4],
//           lo.z[4], hi.z[4], refs[4] (uint bits), unused
//           ref without bit 31: inner node index; ref with bit 31 set: leaf,
//           bits 27..30 = count-1, bits 0..26 = first leaf-ordered triangle;
//           bits 27..31 all set (kInstLeaf): a top-level leaf holding one
//           instance, whose id is in bits 0..26 (two-level scenes only).
//           An empty child slot is a point box at 1e38 with a leaf ref to a
//           degenerate triangle (or to the null instance of a top level).
//   tris  : leaf-ordered triangles, 48 B each = 3 x float4
//           (v0.xyz | global prim id), (v1.xyz | instance id),
//           (v2.xyz | primitive index inside the instance's mesh)
//           Two-level scenes hold each mesh's triangles once, in object
//           space, and the ids are the mesh's own (instance 0, prim id =
//           primitive index): the instance comes from the traversal.
//   insts : 32 B per instance = 2 x float4
//           (albedo.xyz | fuzz for metal, index for glass),
//           (center.xyz = translation of the 3x4 transform | optics type)
//   xinst : two-level scenes only, kInstF4 float4 per instance (+ the null
//           instance of empty top-level slots): world-to-object 3x4 rows
//           (3), object-to-world 3x4 rows (3), (bottom-level root, global id
//           of the instance's first triangle, mesh, 0) as uint bits, unused
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx.h"

namespace rtx {

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr int kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
constexpr uint32_t kMaxLeaf = 8;          // triangles per leaf (<= 16 fits the encoding)
constexpr uint32_t kMaxPrims = kLeafFirstMask;  // 134M triangles per scene
constexpr int kMaxTreeDepth = 64;         // guaranteed by the builder
constexpr int kStackLDS = 32;             // traversal stack entries per lane kept in LDS (default)
constexpr int kMinStackLDS = 16;          // smallest LDS stack a render launch can be tuned to
constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kEmptyRoot = 0xFFFFFFFFu;
constexpr uint32_t kNoChild = 0x7FFFFFFFu;  // empty slot of a 4-wide node (no leaf flag, no valid node)
constexpr int kWideNodeF4 = 8;             // float4 per 4-wide node (128 B)
constexpr uint32_t kInstLeaf = 0xF8000000u; // top-level leaf of one instance: kInstLeaf | instance id
constexpr uint32_t kInstMarker = 0x7FFFFFFEu;  // stack entry below an instance's bottom-level entries
constexpr int kInstF4 = 8;                 // float4 per instance record of a two-level scene (128 B)
#ifndef RTX_TL_FRAC
#define RTX_TL_FRAC 16
#endif
// two-level traversal: instance transitions run once this many 64ths of a
// wave's traversing lanes wait for one (or when the node phase ends)
constexpr int kTlFrac = RTX_TL_FRAC;

// The global address space.  A load through such a pointer compiles to a
// global load.  The batch kernels reach the scene arrays through pointers
// read from the job table, which the compiler cannot place, so it would load
// through them with flat loads; those also count against the LDS wait
// counter, and every LDS stack access would then wait for the node fetches in
// flight.  The batch kernels load the scene arrays through ld4<true>; the
// single-frame kernels' arrays come from kernel arguments, which the compiler
// already places (and the cast costs them 5 % through other code choices).
#define RTX_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const RTX_GLOBAL T* gbl(const T* p) {
    return (const RTX_GLOBAL T*)p;
}
// one float4 through a global pointer (as a native vector: HIP's float4
// class cannot be copied from another address space)
typedef float rtx_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld(const RTX_GLOBAL float4* p) {
    const rtx_f4v v = *(const RTX_GLOBAL rtx_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
// p[0], through the global address space when G
template <bool G>
__device__ __forceinline__ float4 ld4(const float4* p) {
    if constexpr (G) return gld(gbl(p));
    else return *p;
}

// What a traversable handle points at (device memory); kernels get a copy by value.
struct DevScene {
    const float4* nodes;
    const float4* tris;
    const float4* insts;
    const float4* xinst;  // two-level scenes: instance transforms and roots; nullptr: flat scene
    uint32_t root;        // inner node 0, a leaf ref, or kEmptyRoot
    uint32_t num_prims;
    uint32_t num_nodes;   // 4-wide nodes in the node array
    uint32_t num_insts;
};

__host__ __device__ inline uint32_t leaf_ref(uint32_t first, uint32_t count) {
    return kLeafFlag | ((count - 1u) << kLeafCountShift) | first;
}

// ---- small float3 helpers (the V:: operators of optx/v.h:26-53) ------------
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float3 operator-(float3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// optx/v.h:51: 1/len * v
__device__ __forceinline__ float3 unitv(float3 a) { return (1.f / sqrtf(dot(a, a))) * a; }
__device__ __forceinline__ float3 xyz(float4 a) { return f3(a.x, a.y, a.z); }

// ---- Frand48: optx/frand48.h:17-35 -----------------------------------------
// Only the low 48 bits of the 64-bit LCG state reach the output, and they
// depend only on the low 48 bits of the previous state, so the state is
// carried modulo 2^48.  The float is RN(x) * 2^-48 with x < 2^48; RN(x) is
// formed exactly as one correctly rounded add of two exact 24-bit halves.
struct Frand48 {
    uint64_t state;
    __device__ __forceinline__ void init(uint32_t seed) {
        state = seed;
        for (int i = 0; i < 16; i++) (void)next();
    }
    __device__ __forceinline__ float next() {
        state = (0x5DEECE66Dull * state + 0xBull) & 0xFFFFFFFFFFFFull;
        // bits 24..47 by a 32-bit funnel shift of the two state words: as a
        // 64-bit shift the compiler converts the (provably 32-bit) value
        // with the multi-instruction 64-bit integer-to-float sequence
        const uint32_t w0 = (uint32_t)state, w1 = (uint32_t)(state >> 32);
        const float hi = (float)__builtin_amdgcn_alignbit(w1, w0, 24);  // exact
        const float lo = (float)(w0 & 0xFFFFFFu);                        // exact
        // RN(hi + lo * 2^-24) in one rounding, then an exact scale by 2^-24
        return __builtin_fmaf(lo, 0x1p-24f, hi) * 0x1p-24f;
    }
    // util::rnd(min,max): optx/util_gpu.h:28
    __device__ __forceinline__ float range(float lo, float hi) { return lo + next() * (hi - lo); }
};

// optx/v.h:72 (draws x, then y, then z)
__device__ __forceinline__ float3 rnd_in_sphere(Frand48& r) {
    for (;;) {
        const float x = r.range(-1.f, 1.f);
        const float y = r.range(-1.f, 1.f);
        const float z = r.range(-1.f, 1.f);
        const float3 v = f3(x, y, z);
        if (1.f > dot(v, v)) return v;
    }
}
// optx/v.h:75 (draws x, then y)
__device__ __forceinline__ float3 rnd_in_disk(Frand48& r) {
    for (;;) {
        const float x = r.range(-1.f, 1.f);
        const float y = r.range(-1.f, 1.f);
        const float3 v = f3(x, y, 0.f);
        if (1.f > dot(v, v)) return v;
    }
}

// ---- rays and hits ---------------------------------------------------------
struct HitRec {
    float t, u, v;
    uint32_t tri;   // leaf-ordered triangle index, kNoHit on a miss
};

}  // namespace rtx
