// traverse.h -- closest-hit query on the device: the work optixTrace does on
// RT cores in the reference (optx/camera_r.cu:82-99, optx/optics_r.cu:115-132).
//
// Software traversal of the 4-wide BVH (layout in rtx_internal.h) with a
// per-lane stack whose first SLDS entries live in LDS (one column per lane,
// bank-conflict free) and the rest in a global overflow area.  A step pushes
// at most 3 entries per level, so the stack holds at most 3 x (wide-tree
// depth) entries; the launch sizes the overflow area for that.
//
// Triangles use the watertight test of Woop, Benthin and Wald (JCGT 2(1),
// 2013): shared edges and vertices of a closed mesh let no ray through.  The
// edge functions are evaluated without contraction (so the two triangles of
// an edge compute exactly negated values) with an FP64 fallback when one is 0.
// Closest hit: smallest t; an exact tie goes to the lower global primitive id,
// so the result does not depend on traversal order.
//
// Two-level scenes (TL): the top level's leaves are instances.  Entering one
// pushes the world ray (6 words) and a marker on the lane's stack,
// transforms the ray into the instance's object space (the 3x4
// world-to-object rows; the direction is not normalised, so t keeps its
// meaning) and continues at the instance's bottom-level root; popping the
// marker pops the world ray back.  A lane leaves an instance only with no
// leaf parked, and parks no leaf at the top level, so a parked leaf is
// always tested with the ray of its own instance.  Hits keep the instance
// and the primitive index within its mesh; exact ties go to the lower
// (instance, primitive), the order of the global primitive ids.
#pragma once

#include "rtx_internal.h"

namespace rtx {

// slab test pad on the far distance: 1 + 2*gamma(3) (Ize 2013), conservative
constexpr float kBoxPad = 1.0000004f;

struct TRay {
    float3 o, d, idir;
    float tmin;
    // the watertight test's axis order (kx, ky, kz) as a rotation that puts
    // the dominant axis kz last (rz0: kz = x, rz1: kz = y, else kz = z).
    // Woop et al. (2013) also swap kx and ky when d[kz] < 0 to keep the
    // winding; that negates U, V, W, det and T alike and leaves t, u, v bit
    // for bit the same, and nothing here depends on the winding
    bool rz0, rz1;
    float Sx, Sy, Sz;
    // byte offsets, inside a 4-wide node, of the child planes the ray enters
    // through on each axis (lo when the direction component is >= 0, else
    // hi); the exit planes are at offset ^ 16
    uint32_t nx, ny, nz;
    // for slab distances as one FMA per plane, P * idir - c: c = o * idir
    // moved by s = 2^-22 |o * idir| (down for entry planes, up for exit
    // planes), which covers the roundings of o * idir and of c +- s, so
    // an entry distance never comes out above (P - o) * idir rounded
    // once, nor an exit distance below it
    float3 ce, cx;
};


// (a[kx], a[ky], a[kz]) with selects only (no per-lane branches)
__device__ __forceinline__ float3 permute(const TRay& r, float3 a) {
    const float p = r.rz0 ? a.y : (r.rz1 ? a.z : a.x);
    const float q = r.rz0 ? a.z : (r.rz1 ? a.x : a.y);
    const float w = r.rz0 ? a.x : (r.rz1 ? a.y : a.z);
    return f3(p, q, w);
}

__device__ __forceinline__ void tray_init(TRay& r, float3 o, float3 d, float tmin) {
    r.o = o;
    r.d = d;
    r.tmin = tmin;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const float dz = kz == 0 ? d.x : (kz == 1 ? d.y : d.z);
    r.rz0 = kz == 0;
    r.rz1 = kz == 1;
    const float3 dp = permute(r, d);       // (d[kx], d[ky], d[kz])
    r.idir = f3(1.f / (d.x != 0.f ? d.x : 1e-30f), 1.f / (d.y != 0.f ? d.y : 1e-30f),
                1.f / (d.z != 0.f ? d.z : 1e-30f));
    // Sz = 1 / d[kz]: that is idir[kz] unless d is the zero vector
    const float iz = r.rz0 ? r.idir.x : (r.rz1 ? r.idir.y : r.idir.z);
    r.Sz = dz != 0.f ? iz : __builtin_copysignf(__builtin_inff(), dz);
    r.Sx = dp.x * r.Sz;                     // shear by the reciprocal: no division of its own
    r.Sy = dp.y * r.Sz;
    r.nx = r.idir.x >= 0.f ? 0u : 16u;
    r.ny = r.idir.y >= 0.f ? 32u : 48u;
    r.nz = r.idir.z >= 0.f ? 64u : 80u;
    const float3 c = f3(o.x * r.idir.x, o.y * r.idir.y, o.z * r.idir.z);
    const float3 sl = f3(fabsf(c.x) * 0x1p-22f, fabsf(c.y) * 0x1p-22f, fabsf(c.z) * 0x1p-22f);
    r.ce = f3(c.x + sl.x, c.y + sl.y, c.z + sl.z);
    r.cx = f3(c.x - sl.x, c.y - sl.y, c.z - sl.z);
}

// true with t > tmin and barycentrics (u,v) of v1 and v2 on a hit
__device__ __forceinline__ bool watertight(const TRay& r, float3 v0, float3 v1, float3 v2, float& t, float& u,
                                           float& v) {
#pragma clang fp contract(off)
    const float Ax0 = v0.x - r.o.x, Ay0 = v0.y - r.o.y, Az0 = v0.z - r.o.z;
    const float Bx0 = v1.x - r.o.x, By0 = v1.y - r.o.y, Bz0 = v1.z - r.o.z;
    const float Cx0 = v2.x - r.o.x, Cy0 = v2.y - r.o.y, Cz0 = v2.z - r.o.z;
    const float3 A = permute(r, f3(Ax0, Ay0, Az0)), B = permute(r, f3(Bx0, By0, Bz0)),
                 C = permute(r, f3(Cx0, Cy0, Cz0));
    const float Akz = A.z, Bkz = B.z, Ckz = C.z;
    const float Ax = A.x - r.Sx * Akz, Ay = A.y - r.Sy * Akz;
    const float Bx = B.x - r.Sx * Bkz, By = B.y - r.Sy * Bkz;
    const float Cx = C.x - r.Sx * Ckz, Cy = C.y - r.Sy * Ckz;
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
typedef __attribute__((address_space(3))) float4 lds_f4;
typedef __attribute__((address_space(3))) const rtx_f4v lds_f4v;
__device__ __forceinline__ float4 lld4(const char __attribute__((address_space(3)))* p) {
    const rtx_f4v v = *(lds_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

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
    // pushes v[0..N-1] (v[N-1] ends on top) with one range check when all N
    // slots are in LDS
    template <int N>
    __device__ __forceinline__ void push_n(int& sp, const uint32_t (&v)[N]) {
        if (sp + N <= SLDS) {
            lds_u32* p = lds + sp * BLOCK;
#pragma unroll
            for (int i = 0; i < N; i++) p[i * BLOCK] = v[i];
            sp += N;
        } else {
#pragma unroll
            for (int i = 0; i < N; i++) push(sp, v[i]);
        }
    }
    // pops N entries into v[0..N-1] (v[N-1] is the old top), the inverse of push_n
    template <int N>
    __device__ __forceinline__ void pop_n(int& sp, uint32_t (&v)[N]) {
        if (sp <= SLDS) {
            sp -= N;
            const lds_u32* p = lds + sp * BLOCK;
#pragma unroll
            for (int i = 0; i < N; i++) v[i] = p[i * BLOCK];
        } else {
#pragma unroll
            for (int i = N - 1; i >= 0; i--) v[i] = pop(sp);
        }
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
    // One node step: returns the top entry (slot sp-1) and writes the hit
    // children r1..r3 of a sorted 4-wide node above it, farthest lowest:
    // child j (j = 1..nhit-1) goes to slot sp + nhit-1-j.  When every slot
    // the step touches (sp-1 .. sp+2) is in LDS, the writes are unconditional
    // and a child that is not pushed goes to the free slot sp+2 instead.
    __device__ __forceinline__ uint32_t step4(int sp, int nhit, uint32_t r1, uint32_t r2, uint32_t r3) {
        if (sp + 3 <= SLDS) {
            const uint32_t top = lds[(sp > 0 ? sp - 1 : 0) * BLOCK];
            // slot addresses from one base: r1's slot, then one and two
            // slots below it; a child not pushed goes to the free slot sp+2
            // (a base below the column is never selected: p1 is chosen only
            // for hit children, whose slots are >= sp)
            lds_u32* const p1 = lds + (sp + nhit - 2) * BLOCK;
            lds_u32* const fr = lds + (sp + 2) * BLOCK;
            lds_u32* const a3 = nhit > 3 ? p1 - 2 * BLOCK : fr;
            lds_u32* const a2 = nhit > 2 ? p1 - BLOCK : fr;
            lds_u32* const a1 = nhit > 1 ? p1 : fr;
            *a3 = r3;
            *a2 = r2;
            *a1 = r1;
            return top;
        }
        const uint32_t top = peek(sp);
        if (nhit > 1) put(sp + nhit - 2, r1);
        if (nhit > 2) put(sp + nhit - 3, r2);
        if (nhit > 3) put(sp, r3);
        return top;
    }
};

__device__ __forceinline__ float comp4(const float4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}

// compare-exchange of (key, ref) pairs: smaller key first
__device__ __forceinline__ void sort_ce(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
    const bool sw = kb < ka;
    const float k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const uint32_t r0 = sw ? rb : ra, r1 = sw ? ra : rb;
    ka = k0; kb = k1; ra = r0; rb = r1;
}

struct HitFull {
    float t, u, v;
    uint32_t tri;    // leaf-ordered triangle, kNoHit on a miss
    uint32_t prim;   // global primitive id (two-level scenes: primitive index within the instance's mesh)
    uint32_t inst;   // instance of the hit (two-level scenes; flat scenes keep it in the triangle)
};

// Resumable traversal state of one ray: the ray, the next node to visit,
// the stack depth and the best hit so far.
struct Trav {
    TRay r;
    HitFull h;
    uint32_t node;    // next node to visit (kNoChild: none left)
    uint32_t leaf;    // a leaf found but not tested yet (kNoChild: none)
    uint32_t leaf2;   // a second one (set only while leaf is)
    int sp;
    uint32_t inst;    // two-level scenes: the instance the lane is in (kNoHit: top level)
};

__device__ __forceinline__ void trav_reset(Trav& s, const DevScene& sc, float tmax) {
    s.h.t = tmax;
    s.h.u = s.h.v = 0.f;
    s.h.tri = kNoHit;
    s.h.prim = kNoHit;
    s.h.inst = kNoHit;
    s.node = sc.root;
    s.leaf = kNoChild;
    s.leaf2 = kNoChild;
    s.sp = 0;
    s.inst = kNoHit;
}

__device__ __forceinline__ void trav_begin(Trav& s, const DevScene& sc, float3 o, float3 d, float tmin, float tmax) {
    tray_init(s.r, o, d, tmin);
    trav_reset(s, sc, tmax);
}

// row . (p, 1) and row . (v, 0) of a 3x4 transform, one rounding per operation
// (the order of the flattening build, so that a two-level scene's world
// vertices are bit for bit those of the flat scene)
__device__ __forceinline__ float xf_point(const float4& row, float3 p) {
    return ((p.x * row.x + p.y * row.y) + p.z * row.z) + row.w;
}
__device__ __forceinline__ float xf_vector(const float4& row, float3 v) {
    return (v.x * row.x + v.y * row.y) + v.z * row.z;
}
// the same products as fused multiply-adds, for the ray an instance entry
// transforms into object space: nothing has to round like the flat scene
// there, and three FMAs per row replace six (five) operations
__device__ __forceinline__ float xf_point_fma(const float4& row, float3 p) {
    return __builtin_fmaf(p.x, row.x, __builtin_fmaf(p.y, row.y, __builtin_fmaf(p.z, row.z, row.w)));
}
__device__ __forceinline__ float xf_vector_fma(const float4& row, float3 v) {
    return __builtin_fmaf(v.x, row.x, __builtin_fmaf(v.y, row.y, v.z * row.z));
}

// two-level scenes: one instance transition of the lanes that call it.  A
// lane with `leave` set has just popped its instance's marker (and has no
// leaf parked): it pops the world ray and the next entry.  A going lane that
// then stands on an instance leaf enters that instance: it pushes the world
// ray and a marker and continues at the instance's bottom-level root with
// the ray in object space.  The lanes that left, entered or both set their
// new ray up together, in one call.
template <bool G, class Stack>
__device__ __forceinline__ void inst_step(const DevScene& sc, Trav& s, bool going, bool leave, Stack& st) {
    float3 o = s.r.o, d = s.r.d;
    if (leave) {
        uint32_t w[6];
        st.pop_n(s.sp, w);
        o = f3(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]));
        d = f3(__uint_as_float(w[3]), __uint_as_float(w[4]), __uint_as_float(w[5]));
        s.inst = kNoHit;
        s.node = s.sp == 0 ? kNoChild : st.pop(s.sp);
    }
    const bool enter = going && (s.node & kInstLeaf) == kInstLeaf;
    if (enter) {
        const uint32_t id = s.node & kLeafFirstMask;
        const float4* X = sc.xinst + (size_t)kInstF4 * id;
        const float4 r0 = ld4<G>(X), r1 = ld4<G>(X + 1), r2 = ld4<G>(X + 2), m = ld4<G>(X + 6);
        const uint32_t w[7] = {__float_as_uint(o.x), __float_as_uint(o.y), __float_as_uint(o.z), __float_as_uint(d.x),
                               __float_as_uint(d.y), __float_as_uint(d.z), kInstMarker};
        st.push_n(s.sp, w);
        const float3 wo = o, wd = d;
        o = f3(xf_point_fma(r0, wo), xf_point_fma(r1, wo), xf_point_fma(r2, wo));
        d = f3(xf_vector_fma(r0, wd), xf_vector_fma(r1, wd), xf_vector_fma(r2, wd));
        s.inst = id;
        s.node = __float_as_uint(m.x);
    }
    if (leave || enter) tray_init(s.r, o, d, s.r.tmin);
}

// first active lane of the wave (counts wave-level loop iterations in COUNT mode)
__device__ __forceinline__ bool first_active_lane() {
    const uint64_t e = __builtin_amdgcn_read_exec();
    return (threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)e) - 1);
}

// LDS words per wave that the compacted leaf phase uses: the owner marks of
// a round (64), the best key per lane (64 x 64 bit) and the lane holding it
// (64)
constexpr uint32_t kLeafScratchWords = 256;
typedef __attribute__((address_space(3))) uint64_t lds_u64;

// orders the LDS accesses of the wave's lanes around it (the lanes of a wave
// run in lockstep and its LDS operations complete in order, so this only
// keeps the compiler from moving them)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The leaf phase with the work spread over the wave.  A lane's parked
// leaves hold c = c1 + c2 triangles; the wave's (lane, triangle) pairs are
// numbered by an inclusive prefix sum of c and dealt out 64 at a time, one
// per lane whatever lane parked them.  The lane that tests pair p finds its
// owner (the lane whose pairs hold p) by a max-scan over the wave of the
// round positions that the owners mark in LDS, and reads the owner's ray and
// leaves with ds_bpermute.  A hit goes into the owner's LDS slot as the 64-bit minimum
// of (t, primitive id) -- t > tmin >= 0, so the float's bits order like its
// value; the id is global in a flat scene and within the mesh in a
// two-level one, where all of a lane's parked leaves belong to the instance
// it is in -- and the lane whose key won writes its lane number next
// to it; the owner then takes that hit (t, u, v, triangle) under the rule of
// the per-lane loop: closer, or as close with the lower id.  The minimum of
// (t, id) does not depend on the order of the tests, so the closest hit is
// the one the lane would have found testing its own leaves in order.
// Needs all 64 lanes of the wave in the call (the pairs of an inactive lane
// would never be tested), and a ray with tmin >= 0.  A pair is tested only
// if it lies in its owner's own leaves (q < the owner's count), and read
// from the owner's triangle array, so every triangle read is one the owner
// parked.
// max(v, v of the lane that the DPP control CTRL names); lanes outside the
// rows of ROW_MASK, or without a source lane in their row, keep v
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t max_dpp(uint32_t v) {
    const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
    return v > t ? v : t;
}

// v + v of the lane that the DPP control CTRL names (0 where max_dpp keeps v)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t add_dpp(uint32_t v) {
    return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

template <bool G, bool TL, bool COUNT>
__device__ __forceinline__ void leaves_compact(const DevScene& sc, Trav& s, uint32_t f1, uint32_t c1, uint32_t f2,
                                               uint32_t c2, lds_u32* cl, uint32_t* wi) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t c = c1 + c2;
    // inclusive prefix sum of c over the wave (DPP, as the owner scan below)
    uint32_t incl = c;
    incl = add_dpp<0x111, 0xf>(incl);   // row_shr:1
    incl = add_dpp<0x112, 0xf>(incl);   // row_shr:2
    incl = add_dpp<0x114, 0xf>(incl);   // row_shr:4
    incl = add_dpp<0x118, 0xf>(incl);   // row_shr:8
    incl = add_dpp<0x142, 0xa>(incl);   // row_bcast:15 into rows 1 and 3
    incl = add_dpp<0x143, 0xc>(incl);   // row_bcast:31 into rows 2 and 3
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    if (total == 0) return;
    const uint32_t excl = incl - c;
    lds_u32* M = cl;
    lds_u64* K = (lds_u64*)(cl + 64);
    lds_u32* Wn = cl + 192;
    K[lane] = ~0ull;
    // what a helper lane needs of the owner: its ray (object space inside an
    // instance), the first pair's number, and its leaves
    const uint32_t meta = excl | (c1 << 16) | (c << 21) | (s.r.rz0 ? 1u << 30 : 0u) | (s.r.rz1 ? 1u << 31 : 0u);
    wave_sync();
    for (uint32_t base = 0; base < total; base += 64u) {
        if (COUNT && lane == 0) wi[1]++;
        const uint32_t p = base + lane;
        const bool has = p < total;
        // owner: the lane whose pairs [excl, incl) hold p.  Each lane whose
        // pairs meet this round marks the round position of its first one
        // with its lane number; the owner of position j is then the largest
        // mark at or before j (lane numbers grow with the pair numbers), an
        // inclusive max-scan over the wave (DPP: shifts within the 16-lane
        // rows, then broadcasts of the row ends to the rows above)
        M[lane] = 0u;
        wave_sync();
        if (c != 0u && incl > base && excl < base + 64u) M[(excl > base ? excl : base) - base] = lane;
        wave_sync();
        uint32_t i = M[lane];
        i = max_dpp<0x111, 0xf>(i);   // row_shr:1
        i = max_dpp<0x112, 0xf>(i);   // row_shr:2
        i = max_dpp<0x114, 0xf>(i);   // row_shr:4
        i = max_dpp<0x118, 0xf>(i);   // row_shr:8
        i = max_dpp<0x142, 0xa>(i);   // row_bcast:15 into rows 1 and 3
        i = max_dpp<0x143, 0xc>(i);   // row_bcast:31 into rows 2 and 3
        const int ad = (int)(i << 2);
        auto bp = [&](uint32_t v) -> uint32_t { return (uint32_t)__builtin_amdgcn_ds_bpermute(ad, (int)v); };
        auto bpf = [&](float v) -> float { return __uint_as_float(bp(__float_as_uint(v))); };
        TRay hr;
        hr.o = f3(bpf(s.r.o.x), bpf(s.r.o.y), bpf(s.r.o.z));
        hr.Sx = bpf(s.r.Sx);
        hr.Sy = bpf(s.r.Sy);
        hr.Sz = bpf(s.r.Sz);
        hr.tmin = bpf(s.r.tmin);
        const uint32_t om = bp(meta), of1 = bp(f1), of2 = bp(f2);
        hr.rz0 = (om >> 30) & 1u;
        hr.rz1 = (om >> 31) & 1u;
        const uint32_t q = p - (om & 0xFFFFu), oc1 = (om >> 16) & 31u, oc = (om >> 21) & 63u;
        const uint32_t k = q < oc1 ? of1 + q : of2 + (q - oc1);
        float tt = 0.f, uu = 0.f, vv = 0.f;
        bool hit = false;
        uint64_t key = ~0ull;
        // the owner's triangle array: in a batch (G) each lane renders its
        // own job, so the owner's scene need not be this lane's
        const char* tb = (const char*)sc.tris;
        if (G) {
            const uint64_t pt = (uint64_t)sc.tris;
            tb = (const char*)(((uint64_t)bp((uint32_t)(pt >> 32)) << 32) | (uint64_t)bp((uint32_t)pt));
        }
        if (has && q < oc) {
            const uint32_t to = k * 48u;
            typedef const float4 cf4t;
            const float4 t0 = ld4<G>((cf4t*)(tb + to)), t1 = ld4<G>((cf4t*)(tb + (to + 16u))),
                         t2 = ld4<G>((cf4t*)(tb + (to + 32u)));
            hit = watertight(hr, xyz(t0), xyz(t1), xyz(t2), tt, uu, vv);
            if (hit) {
                key = ((uint64_t)__float_as_uint(tt) << 32) | (uint64_t)__float_as_uint(t0.w);
                __hip_atomic_fetch_min(K + i, key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        }
        wave_sync();
        if (hit && K[i] == key) Wn[i] = lane;
        wave_sync();
        // the owners of this round's pairs take their best hit
        const uint64_t kk = K[lane];
        const int wa = (int)((Wn[lane] & 63u) << 2);
        const float bu = __uint_as_float((uint32_t)__builtin_amdgcn_ds_bpermute(wa, (int)__float_as_uint(uu)));
        const float bv = __uint_as_float((uint32_t)__builtin_amdgcn_ds_bpermute(wa, (int)__float_as_uint(vv)));
        const uint32_t bk = (uint32_t)__builtin_amdgcn_ds_bpermute(wa, (int)k);
        // (a slot holds a key only if its lane owned a pair of this round
        // that hit: every slot is reset when it is taken)
        if (kk != ~0ull) {
            const float t = __uint_as_float((uint32_t)(kk >> 32));
            const uint32_t prim = (uint32_t)kk;
            const bool lower = TL ? (s.inst < s.h.inst || (s.inst == s.h.inst && prim < s.h.prim)) : prim < s.h.prim;
            if (t < s.h.t || (t == s.h.t && lower)) {
                s.h.t = t;
                s.h.u = bu;
                s.h.v = bv;
                s.h.tri = bk;
                s.h.prim = prim;
                if (TL) s.h.inst = s.inst;
            }
            K[lane] = ~0ull;
        }
        wave_sync();
    }
}

// Advances the traversal of the lanes with `going` set.  A lane clears
// `going` when its ray is done (closest hit final in s.h).
//
// Each round has a node phase and a leaf phase.  The node phase steps the
// lanes that sit on an inner node; a lane that reaches a leaf parks it in
// one of two leaf slots and goes on with the next node from its stack (Aila
// and Laine 2009, "speculative traversal"), and waits only when it reaches a
// third leaf or runs out of nodes.  The phase ends when fewer than
// inner_frac/64 of the going lanes can still step; the leaf phase then tests
// the parked leaves of all lanes together.  inner_frac = 1 runs the node
// phase until no lane can step.  The function
// returns once no lane is going, or -- with `batch` < 64 -- as soon as at
// least `batch` of the wave's `busy` lanes have finished, so that the caller
// can shade them while the others keep their state and resume later.
// COUNT adds node-visit and triangle-test counts (statistics launches only).
// FMA selects the slab form: one FMA per plane with the per-ray offsets
// (fewer instructions), or (P - o) * idir (six fewer registers, for kernels
// that are short of them).
// TOPN > 0: the first TOPN nodes of the node array (the top levels: the
// array is breadth-first) are read from a copy in LDS at `top` instead of
// global memory.
// CL: the leaf phase spreads the parked leaves' triangles over the wave
// (leaves_compact, with the wave's kLeafScratchWords of LDS at `cl`);
// otherwise each lane tests its own leaves.
template <int BLOCK, bool COUNT = false, bool TL = false, bool FMA = true, bool G = false, int TOPN = 0,
          bool CL = false, class Stack>
__device__ __forceinline__ void trav_run(const DevScene& sc, Trav& s, bool& going, bool busy, int batch,
                                         int inner_frac, Stack& st, uint32_t* nn = nullptr,
                                         uint32_t* nt = nullptr, uint32_t* wi = nullptr,
                                         const lds_f4* top = nullptr, lds_u32* cl = nullptr) {
    if (going && s.node == kEmptyRoot) going = false;
    for (;;) {
        // node phase (no lane starts or stops going in it)
        const uint64_t m_going = __ballot(going);
        const uint32_t n_going = (uint32_t)__popcll(m_going);
        for (;;) {
            if (TL) {
                // Instance transitions: leave a finished instance (not while
                // one of its leaves is parked), then enter the instance the
                // lane stands on, if any.  A transition costs several node
                // steps of instructions, so a lane that needs one waits
                // (without stepping) until kTlFrac/64 of the going lanes do,
                // or until the round's transition step after the leaf phase.
                const bool leave = going && s.node == kInstMarker && s.leaf == kNoChild;
                const bool enter = going && (s.node & kInstLeaf) == kInstLeaf;
                const uint64_t need = __ballot(leave || enter);
                if (need && (uint32_t)__popcll(need) * 64u >= (uint32_t)kTlFrac * n_going)
                    inst_step<G>(sc, s, going, leave, st);
            }
            // a lane that reached a leaf parks it (two slots) and goes on
            // with the next node from its stack, so it keeps stepping while
            // the others do; the parked leaves are tested together below
            // (an instance leaf waits for its transition instead)
            if (going && (s.node & kLeafFlag) && (!TL || (s.node & kInstLeaf) != kInstLeaf) &&
                (s.leaf == kNoChild || s.leaf2 == kNoChild)) {
                if (s.leaf == kNoChild) s.leaf = s.node;
                else s.leaf2 = s.node;
                s.node = s.sp == 0 ? kNoChild : st.pop(s.sp);
            }
            // no leaf flag, not kNoChild (nor the instance marker): one unsigned
            // compare, balloted as it is and masked with the going lanes
            const bool below = s.node < (TL ? kInstMarker : kNoChild);
            const bool inner = going && below;
            const uint32_t n_inner = (uint32_t)__popcll(__ballot(below) & m_going);
            if (n_inner == 0) break;
            if (n_inner * 64u < (uint32_t)inner_frac * n_going) break;
            // the node step as a block, not behind a `continue`: every exit of
            // this loop is then wave-uniform (the two breaks above), and the
            // loop's control stays in scalar registers (2-3 % faster frames)
            if (inner) {
                if (COUNT) {
                    (*nn)++;
                    if (first_active_lane()) wi[0]++;
                }
                // one 4-wide node: six float4 of child bounds (SoA: lo.x, hi.x,
                // lo.y, hi.y, lo.z, hi.z) + one of refs.  The ray loads each
                // axis's entry and exit planes by its direction signs, so the
                // slab test needs no min/max per axis.
                // (32-bit byte offsets: the builder keeps the node array below 4 GiB)
                typedef const float4 cf4;
                const char* nb = (const char*)sc.nodes;
                const uint32_t no = s.node * (16u * kWideNodeF4);
                const TRay& r = s.r;
                float4 ex, fx, ey, fy, ez, fz, rf;
                if (TOPN > 0 && s.node < (uint32_t)TOPN) {
                    typedef const char __attribute__((address_space(3))) lchar;
                    lchar* lb = (lchar*)top;
                    ex = lld4(lb + (no + r.nx));
                    fx = lld4(lb + (no + (r.nx ^ 16u)));
                    ey = lld4(lb + (no + r.ny));
                    fy = lld4(lb + (no + (r.ny ^ 16u)));
                    ez = lld4(lb + (no + r.nz));
                    fz = lld4(lb + (no + (r.nz ^ 16u)));
                    rf = lld4(lb + (no + 96u));
                } else {
                    ex = ld4<G>((cf4*)(nb + (no + r.nx)));
                    fx = ld4<G>((cf4*)(nb + (no + (r.nx ^ 16u))));
                    ey = ld4<G>((cf4*)(nb + (no + r.ny)));
                    fy = ld4<G>((cf4*)(nb + (no + (r.ny ^ 16u))));
                    ez = ld4<G>((cf4*)(nb + (no + r.nz)));
                    fz = ld4<G>((cf4*)(nb + (no + (r.nz ^ 16u))));
                    rf = ld4<G>((cf4*)(nb + (no + 96u)));
                }
                float key[4];
                uint32_t ref[4];
                int nhit = 0;
    #pragma unroll
                for (int j = 0; j < 4; j++) {
                    float x0, x1, y0, y1, z0, z1;
                    if (FMA) {
                        x0 = __builtin_fmaf(comp4(ex, j), r.idir.x, -r.ce.x);
                        x1 = __builtin_fmaf(comp4(fx, j), r.idir.x, -r.cx.x);
                        y0 = __builtin_fmaf(comp4(ey, j), r.idir.y, -r.ce.y);
                        y1 = __builtin_fmaf(comp4(fy, j), r.idir.y, -r.cx.y);
                        z0 = __builtin_fmaf(comp4(ez, j), r.idir.z, -r.ce.z);
                        z1 = __builtin_fmaf(comp4(fz, j), r.idir.z, -r.cx.z);
                    } else {
                        x0 = (comp4(ex, j) - r.o.x) * r.idir.x;
                        x1 = (comp4(fx, j) - r.o.x) * r.idir.x;
                        y0 = (comp4(ey, j) - r.o.y) * r.idir.y;
                        y1 = (comp4(fy, j) - r.o.y) * r.idir.y;
                        z0 = (comp4(ez, j) - r.o.z) * r.idir.z;
                        z1 = (comp4(fz, j) - r.o.z) * r.idir.z;
                    }
                    const float tn = fmaxf(fmaxf(x0, y0), fmaxf(z0, r.tmin));
                    const float tf = fminf(fminf(x1, y1), fminf(z1, s.h.t));
                    // empty slots hold a point box at 1e38 that no slab test accepts
                    const uint32_t c = __float_as_uint(comp4(rf, j));
                    const bool hit = tn <= tf * kBoxPad;
                    key[j] = hit ? tn : __uint_as_float(0x7F800000u);
                    ref[j] = c;
                    nhit += hit ? 1 : 0;
                }
                // ascending by entry distance (misses last): 5 compare-exchanges
                sort_ce(key[0], ref[0], key[1], ref[1]);
                sort_ce(key[2], ref[2], key[3], ref[3]);
                sort_ce(key[0], ref[0], key[2], ref[2]);
                sort_ce(key[1], ref[1], key[3], ref[3]);
                sort_ce(key[1], ref[1], key[2], ref[2]);
                // visit the nearest; push the others farthest first (slot
                // sp + nhit-1-j for child j), so the next nearest is on top.
                // With no hit the top of the stack is next (read before any
                // write: the writes go to slots >= sp).
                const uint32_t top = st.step4(s.sp, nhit, ref[1], ref[2], ref[3]);
                if (nhit == 0) {
                    s.node = s.sp == 0 ? kNoChild : top;
                    s.sp = s.sp > 0 ? s.sp - 1 : 0;
                } else {
                    s.sp += nhit - 1;
                    s.node = ref[0];
                }
            }
        }
        // leaf phase (compacted when the whole wave is here)
        if (CL && __builtin_amdgcn_read_exec() == ~0ull) {
            uint32_t f1 = 0, c1 = 0, f2 = 0, c2 = 0;
            if (going && s.leaf != kNoChild) {
                f1 = s.leaf & kLeafFirstMask;
                c1 = ((s.leaf >> kLeafCountShift) & 15u) + 1u;
                if (s.leaf2 != kNoChild) {
                    f2 = s.leaf2 & kLeafFirstMask;
                    c2 = ((s.leaf2 >> kLeafCountShift) & 15u) + 1u;
                }
                s.leaf = kNoChild;
                s.leaf2 = kNoChild;
            }
            if (COUNT) *nt += c1 + c2;
            leaves_compact<G, TL, COUNT>(sc, s, f1, c1, f2, c2, cl, wi);
        } else if (going && s.leaf != kNoChild) {
            const uint32_t lf = s.leaf;
            s.leaf = kNoChild;
            const uint32_t first = lf & kLeafFirstMask;
            const uint32_t cnt = ((lf >> kLeafCountShift) & 15u) + 1u;
            const uint32_t lf2 = s.leaf2;
            s.leaf2 = kNoChild;
            const uint32_t first2 = lf2 & kLeafFirstMask;
            const uint32_t cnt2 = lf2 == kNoChild ? 0u : ((lf2 >> kLeafCountShift) & 15u) + 1u;
            if (COUNT) *nt += cnt + cnt2;
            // the triangles of both parked leaves in one loop
            for (uint32_t q = 0; q < cnt + cnt2; q++) {
                const uint32_t k = q < cnt ? first + q : first2 + (q - cnt);
                if (COUNT && first_active_lane()) wi[1]++;
                // (32-bit byte offsets: the builder keeps the triangle array below 4 GiB)
                const char* tb = (const char*)sc.tris;
                const uint32_t to = k * 48u;
                typedef const float4 cf4t;
                const float4 t0 = ld4<G>((cf4t*)(tb + to)), t1 = ld4<G>((cf4t*)(tb + (to + 16u))),
                             t2 = ld4<G>((cf4t*)(tb + (to + 32u)));
                float tt, uu, vv;
                if (watertight(s.r, xyz(t0), xyz(t1), xyz(t2), tt, uu, vv)) {
                    const uint32_t prim = __float_as_uint(t0.w);
                    const bool lower = TL ? (s.inst < s.h.inst || (s.inst == s.h.inst && prim < s.h.prim))
                                          : prim < s.h.prim;
                    if (tt < s.h.t || (tt == s.h.t && lower)) {
                        s.h.t = tt;
                        s.h.u = uu;
                        s.h.v = vv;
                        s.h.tri = k;
                        s.h.prim = prim;
                        if (TL) s.h.inst = s.inst;
                    }
                }
            }
        }
        if (TL) {
            // the round's transition step: every lane waiting for one (no
            // leaf is parked after the leaf phase)
            const bool leave = going && s.node == kInstMarker;
            const bool enter = going && (s.node & kInstLeaf) == kInstLeaf;
            if (__ballot(leave || enter)) inst_step<G>(sc, s, going, leave, st);
        }
        if (going && s.node == kNoChild && s.leaf == kNoChild) going = false;
        const uint64_t g = __ballot(going);
        if (g == 0) return;
        if (batch < 64 && __popcll(__ballot(busy && !going)) >= batch) return;
    }
}

// closest hit, run to completion (query kernels and the picker)
template <int BLOCK, bool TL = false, class Stack>
__device__ __forceinline__ void traverse(const DevScene& sc, const TRay& r, float tmax, HitFull& h, Stack& st) {
    Trav s;
    s.r = r;
    trav_reset(s, sc, tmax);
    bool going = true;
    trav_run<BLOCK, false, TL>(sc, s, going, true, 64, 32, st);
    h = s.h;
}

}  // namespace rtx
