// traverse.h -- closest-hit query on the device: the work optixTrace does on
// RT cores in the reference (optx/camera_r.cu:82-99, optx/optics_r.cu:115-132).
//
// Software traversal of the two-child BVH (layout in rtx_internal.h) with a
// per-lane stack whose first kStackLDS entries live in LDS (one column per
// lane, bank-conflict free) and the rest in a global overflow area.  The
// builder bounds the tree depth by kMaxTreeDepth, so the stack never holds
// more than kMaxTreeDepth - 1 entries.
//
// Triangles use the watertight test of Woop, Benthin and Wald (JCGT 2(1),
// 2013): shared edges and vertices of a closed mesh let no ray through.  The
// edge functions are evaluated without contraction (so the two triangles of
// an edge compute exactly negated values) with an FP64 fallback when one is 0.
// Closest hit: smallest t; an exact tie goes to the lower global primitive id,
// so the result does not depend on traversal order.
#pragma once

#include "rtx_internal.h"

namespace rtx {

// slab test pad on the far distance: 1 + 2*gamma(3) (Ize 2013), conservative
constexpr float kBoxPad = 1.0000004f;

struct TRay {
    float3 o, d, idir;
    float tmin;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float comp(const float3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ void tray_init(TRay& r, float3 o, float3 d, float tmin) {
    r.o = o;
    r.d = d;
    r.tmin = tmin;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dz = comp(d, kz);
    if (dz < 0.f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.f / dz;
    r.idir = f3(1.f / (d.x != 0.f ? d.x : 1e-30f), 1.f / (d.y != 0.f ? d.y : 1e-30f),
                1.f / (d.z != 0.f ? d.z : 1e-30f));
}

// true with t > tmin and barycentrics (u,v) of v1 and v2 on a hit
__device__ __forceinline__ bool watertight(const TRay& r, float3 v0, float3 v1, float3 v2, float& t, float& u,
                                           float& v) {
#pragma clang fp contract(off)
    const float Ax0 = v0.x - r.o.x, Ay0 = v0.y - r.o.y, Az0 = v0.z - r.o.z;
    const float Bx0 = v1.x - r.o.x, By0 = v1.y - r.o.y, Bz0 = v1.z - r.o.z;
    const float Cx0 = v2.x - r.o.x, Cy0 = v2.y - r.o.y, Cz0 = v2.z - r.o.z;
    const float3 A = f3(Ax0, Ay0, Az0), B = f3(Bx0, By0, Bz0), C = f3(Cx0, Cy0, Cz0);
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Akz, Ay = comp(A, r.ky) - r.Sy * Akz;
    const float Bx = comp(B, r.kx) - r.Sx * Bkz, By = comp(B, r.ky) - r.Sy * Bkz;
    const float Cx = comp(C, r.kx) - r.Sx * Ckz, Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.f / det;
    t = T * rcp;
    if (!(t > r.tmin)) return false;
    u = V * rcp;
    v = W * rcp;
    return true;
}

// LDS addresses as such (address space 3), so that stack accesses compile to
// ds_read/ds_write instead of flat loads through a generic pointer
typedef __attribute__((address_space(3))) uint32_t lds_u32;

template <int BLOCK, int SLDS = kStackLDS>
struct LaneStack {
    lds_u32* lds;      // LDS column of this lane: entry i at lds[i * BLOCK]
    uint32_t* glob;    // overflow column: entry i at glob[(i - SLDS) * gstride]
    uint32_t gstride;
    __device__ __forceinline__ void push(int& sp, uint32_t v) {
        if (sp < SLDS) lds[sp * BLOCK] = v;
        else glob[(size_t)(sp - SLDS) * gstride] = v;
        sp++;
    }
    __device__ __forceinline__ uint32_t pop(int& sp) {
        sp--;
        uint32_t v;
        if (sp < SLDS) v = lds[sp * BLOCK];
        else v = glob[(size_t)(sp - SLDS) * gstride];
        return v;
    }
    // writes slot sp (the free slot above the top) without pushing
    __device__ __forceinline__ void put(int sp, uint32_t v) {
        if (sp < SLDS) lds[sp * BLOCK] = v;
        else glob[(size_t)(sp - SLDS) * gstride] = v;
    }
    // reads the top entry (slot sp-1; slot 0 when empty, value unused then)
    __device__ __forceinline__ uint32_t peek(int sp) {
        const int i = sp > 0 ? sp - 1 : 0;
        if (i < SLDS) return lds[i * BLOCK];
        return glob[(size_t)(i - SLDS) * gstride];
    }
};

struct HitFull {
    float t, u, v;
    uint32_t tri;    // leaf-ordered triangle, kNoHit on a miss
    uint32_t prim;   // global primitive id
};

// Resumable traversal state of one ray: the ray, the next node to visit,
// the stack depth and the best hit so far.
struct Trav {
    TRay r;
    HitFull h;
    uint32_t node;
    int sp;
};

__device__ __forceinline__ void trav_begin(Trav& s, const DevScene& sc, float3 o, float3 d, float tmin, float tmax) {
    tray_init(s.r, o, d, tmin);
    s.h.t = tmax;
    s.h.u = s.h.v = 0.f;
    s.h.tri = kNoHit;
    s.h.prim = kNoHit;
    s.node = sc.root;
    s.sp = 0;
}

// first active lane of the wave (counts wave-level loop iterations in COUNT mode)
__device__ __forceinline__ bool first_active_lane() {
    const uint64_t e = __builtin_amdgcn_read_exec();
    return (threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)e) - 1);
}

// Advances the traversal of the lanes with `going` set.  A lane clears
// `going` when its ray is done (closest hit final in s.h).
//
// Each round has a node phase and a leaf phase.  The node phase steps the
// lanes that sit on an inner node and ends when fewer than inner_frac/64 of
// the going lanes still do (lanes that reached a leaf wait for it); the leaf
// phase tests the triangles of every lane on a leaf.  inner_frac = 1 gives
// the classic while-while loop (all lanes reach a leaf first).  The function
// returns once no lane is going, or -- with `batch` < 64 -- as soon as at
// least `batch` of the wave's `busy` lanes have finished, so that the caller
// can shade them while the others keep their state and resume later.
// COUNT adds node-visit and triangle-test counts (statistics launches only).
template <int BLOCK, bool COUNT = false, class Stack>
__device__ __forceinline__ void trav_run(const DevScene& sc, Trav& s, bool& going, bool busy, int batch,
                                         int inner_frac, Stack& st, uint32_t* nn = nullptr,
                                         uint32_t* nt = nullptr, uint32_t* wi = nullptr) {
    if (going && s.node == kEmptyRoot) going = false;
    for (;;) {
        // node phase
        for (;;) {
            const bool inner = going && !(s.node & kLeafFlag);
            const uint32_t n_inner = (uint32_t)__popcll(__ballot(inner));
            if (n_inner == 0) break;
            if (n_inner * 64u < (uint32_t)inner_frac * (uint32_t)__popcll(__ballot(going))) break;
            if (!inner) continue;
            if (COUNT) {
                (*nn)++;
                if (first_active_lane()) wi[0]++;
            }
            const float4* np = sc.nodes + 4 * (size_t)s.node;
            const float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
            const TRay& r = s.r;
            const float a0 = (n0.x - r.o.x) * r.idir.x, a1 = (n0.y - r.o.x) * r.idir.x;
            const float a2 = (n0.z - r.o.y) * r.idir.y, a3 = (n0.w - r.o.y) * r.idir.y;
            const float a4 = (n2.x - r.o.z) * r.idir.z, a5 = (n2.y - r.o.z) * r.idir.z;
            const float b0 = (n1.x - r.o.x) * r.idir.x, b1 = (n1.y - r.o.x) * r.idir.x;
            const float b2 = (n1.z - r.o.y) * r.idir.y, b3 = (n1.w - r.o.y) * r.idir.y;
            const float b4 = (n2.z - r.o.z) * r.idir.z, b5 = (n2.w - r.o.z) * r.idir.z;
            const float c0min = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r.tmin));
            const float c0max = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), s.h.t));
            const float c1min = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r.tmin));
            const float c1max = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), s.h.t));
            const bool hit0 = c0min <= c0max * kBoxPad;
            const bool hit1 = c1min <= c1max * kBoxPad;
            // branch-free child order and stack update: the far child is
            // written to the free slot every step and kept only if both
            // children are hit; the top slot is read every step and used
            // only if neither is
            const bool both = hit0 && hit1;
            const bool take1 = both ? (c1min < c0min) : hit1;
            const uint32_t ca = __float_as_uint(n3.x), cb = __float_as_uint(n3.y);
            const uint32_t nearc = take1 ? cb : ca, farc = take1 ? ca : cb;
            st.put(s.sp, farc);
            const uint32_t top = st.peek(s.sp);
            const bool any = hit0 || hit1;
            s.sp += both ? 1 : 0;
            if (!any) {
                if (s.sp == 0) going = false;
                s.sp = s.sp > 0 ? s.sp - 1 : 0;
            }
            s.node = any ? nearc : top;
        }
        // leaf phase
        if (going && (s.node & kLeafFlag)) {
            const uint32_t first = s.node & kLeafFirstMask;
            const uint32_t cnt = ((s.node >> kLeafCountShift) & 15u) + 1u;
            if (COUNT) *nt += cnt;
            for (uint32_t k = first; k < first + cnt; k++) {
                if (COUNT && first_active_lane()) wi[1]++;
                const float4* tp = sc.tris + 3 * (size_t)k;
                const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
                float tt, uu, vv;
                if (watertight(s.r, xyz(t0), xyz(t1), xyz(t2), tt, uu, vv)) {
                    const uint32_t prim = __float_as_uint(t0.w);
                    if (tt < s.h.t || (tt == s.h.t && prim < s.h.prim)) {
                        s.h.t = tt;
                        s.h.u = uu;
                        s.h.v = vv;
                        s.h.tri = k;
                        s.h.prim = prim;
                    }
                }
            }
            if (s.sp == 0) going = false;
            else s.node = st.pop(s.sp);
        }
        const uint64_t g = __ballot(going);
        if (g == 0) return;
        if (batch < 64 && __popcll(__ballot(busy && !going)) >= batch) return;
    }
}

// closest hit, run to completion (query kernels and the picker)
template <int BLOCK, bool COUNT = false, class Stack>
__device__ __forceinline__ void traverse(const DevScene& sc, const TRay& r, float tmax, HitFull& h,
                                         Stack& st, uint32_t* nn = nullptr, uint32_t* nt = nullptr) {
    Trav s;
    s.r = r;
    s.h.t = tmax;
    s.h.u = s.h.v = 0.f;
    s.h.tri = kNoHit;
    s.h.prim = kNoHit;
    s.node = sc.root;
    s.sp = 0;
    bool going = true;
    trav_run<BLOCK, COUNT>(sc, s, going, true, 64, 32, st, nn, nt);
    h = s.h;
}

}  // namespace rtx
