// bvh_build.hip -- on-device acceleration structure build.
//
// Replaces the OptiX GAS/IAS builds of optx/scene.cxx:38-273 (optixAccelBuild)
// with: instance flattening to world space, a top-down binned SAH BVH built
// level by level on the device (Wald 2007, "On fast construction of SAH-based
// bounding volume hierarchies"), triangle reordering into leaf order, and a
// collapse into the 128-byte 4-wide node layout the traversal reads.  The
// same builder makes the levels of a two-level scene: one tree per mesh over
// its object-space triangles and one over the instances' world boxes (one
// instance per leaf); tlas.hip assembles and refits them.
//
// Determinism: every level is processed in task order, child nodes and tasks
// are placed by an exclusive scan (no atomics decide positions) and the
// partition is stable, so the same input always gives the same BVH.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "bvh_build.h"

namespace rtx {

namespace {

#ifndef RTX_SAH_BINS
#define RTX_SAH_BINS 64
#endif
constexpr int kBins = RTX_SAH_BINS;
constexpr int kBuildBlock = 256;
constexpr uint32_t kSmallTask = 256;   // primitives per task below which a level runs a wave per task
constexpr uint32_t kWideTasks = 512;      // levels of at most this many tasks, averaging ...
constexpr uint32_t kWideMinPrims = 8192;  // ... at least this many primitives, spread each task
constexpr uint32_t kWideBlocks = 1024;    // over kWideBlocks / tasks chunks, ...
constexpr uint32_t kWideMinChunk = 2048;  // ... of this many primitives at least (on average)
constexpr uint32_t kRootBlocks = 2048;                   // blocks of the root bounds reduction at most
constexpr int kScanBlock = 1024;
// SAH constants (cost of one node visit vs one triangle test)
// 1: the 4-wide tree is the SAH-optimal grouping of the binary one; 0: each
// wide node opens its largest-area inner child until it has 4 children
#ifndef RTX_DP_COLLAPSE
#define RTX_DP_COLLAPSE 1
#endif
#ifndef RTX_SAH_TRAV
#define RTX_SAH_TRAV 1.0f
#endif
constexpr float kCostTrav = RTX_SAH_TRAV;
constexpr float kCostIsect = 1.0f;
constexpr float kFarAway = 1e38f;   // coordinate of the empty-slot boxes and the degenerate triangle

struct BuildTask {
    uint32_t node, begin, end, depth;
    float4 clo, chi;   // centroid bounds of the task's primitives
};

struct SplitOut {
    uint32_t kind;     // 0 leaf, 1 split
    uint32_t nleft;
    uint32_t pad0, pad1;
    float4 clo_l, chi_l, clo_r, chi_r;
};

// binary node during the build: lo.w/hi.w carry the links as uint bits
//   inner: lo.w = left child (right = left + 1), hi.w = 0
//   leaf : lo.w = first leaf-ordered primitive, hi.w = count (>= 1)
struct BNode {
    float4 lo, hi;
};

struct BuildStats {
    uint32_t leaves, max_depth, max_leaf, pad;
    double sah;
};

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ float area3(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    if (!(dx >= 0.f) || !(dy >= 0.f) || !(dz >= 0.f)) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ float comp4(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ int bin_of(float c, float lo, float scale) {
    int b = (int)((c - lo) * scale);
    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

// ---------------------------------------------------------------------------
// flatten: one thread per world-space triangle (instance-major order)
// world = R v + t evaluated as ((v.x*t0 + v.y*t1) + v.z*t2) + t3, one
// rounding per operation (optx/optics_r.cu:61-75 with contraction ruled out)
// ---------------------------------------------------------------------------
__global__ void k_flatten(const InstBuild* __restrict__ insts, uint32_t ninst,
                          const MeshRef* __restrict__ meshes, uint32_t N,
                          float4* __restrict__ tri_world, float4* __restrict__ plo,
                          float4* __restrict__ phi) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    uint32_t lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (insts[mid].tri_base <= p) lo = mid; else hi = mid - 1;
    }
    const InstBuild& in = insts[lo];
    const uint32_t j = p - in.tri_base;
    const MeshRef m = meshes[in.mesh];
    float3 w[3];
    for (int c = 0; c < 3; c++) {
        const uint32_t vi = m.ices[3 * j + c];
        const float vx = m.vces[3 * vi + 0], vy = m.vces[3 * vi + 1], vz = m.vces[3 * vi + 2];
        float r[3];
        for (int k = 0; k < 3; k++) {
            float acc = __fmul_rn(vx, in.xf[4 * k + 0]);
            acc = __fadd_rn(acc, __fmul_rn(vy, in.xf[4 * k + 1]));
            acc = __fadd_rn(acc, __fmul_rn(vz, in.xf[4 * k + 2]));
            r[k] = __fadd_rn(acc, in.xf[4 * k + 3]);
        }
        w[c] = make_float3(r[0], r[1], r[2]);
    }
    tri_world[3 * (size_t)p + 0] = make_float4(w[0].x, w[0].y, w[0].z, __uint_as_float(p));
    tri_world[3 * (size_t)p + 1] = make_float4(w[1].x, w[1].y, w[1].z, __uint_as_float(lo));
    tri_world[3 * (size_t)p + 2] = make_float4(w[2].x, w[2].y, w[2].z, __uint_as_float(j));
    plo[p] = make_float4(fminf(fminf(w[0].x, w[1].x), w[2].x), fminf(fminf(w[0].y, w[1].y), w[2].y),
                         fminf(fminf(w[0].z, w[1].z), w[2].z), 0.f);
    phi[p] = make_float4(fmaxf(fmaxf(w[0].x, w[1].x), w[2].x), fmaxf(fmaxf(w[0].y, w[1].y), w[2].y),
                         fmaxf(fmaxf(w[0].z, w[1].z), w[2].z), 0.f);
}

// root task: centroid bounds of all primitives as ordered ints (minima of x,
// y, z, then maxima), reduced per block and merged across blocks by atomics;
// min and max are exact, so the result does not depend on their order
__global__ void k_root_init(uint32_t* __restrict__ rb) {
    if (threadIdx.x < 6) rb[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
}
__global__ __launch_bounds__(256) void k_root_bounds(const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                     uint32_t N, uint32_t* __restrict__ rb) {
    __shared__ uint32_t s_b[6];
    if (threadIdx.x < 6) s_b[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const float4 a = plo[p], b = phi[p];
        const float c[3] = {.5f * (a.x + b.x), .5f * (a.y + b.y), .5f * (a.z + b.z)};
        for (int k = 0; k < 3; k++) {
            lo[k] = min(lo[k], f2o(c[k]));
            hi[k] = max(hi[k], f2o(c[k]));
        }
    }
    for (int k = 0; k < 3; k++) { atomicMin(&s_b[k], lo[k]); atomicMax(&s_b[3 + k], hi[k]); }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&rb[threadIdx.x], s_b[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&rb[threadIdx.x], s_b[threadIdx.x]);
}
__global__ void k_root_task(const uint32_t* __restrict__ rb, uint32_t N, BuildTask* __restrict__ task) {
    if (threadIdx.x != 0) return;
    BuildTask t;
    t.node = 0; t.begin = 0; t.end = N; t.depth = 0;
    t.clo = make_float4(o2f(rb[0]), o2f(rb[1]), o2f(rb[2]), 0.f);
    t.chi = make_float4(o2f(rb[3]), o2f(rb[4]), o2f(rb[5]), 0.f);
    *task = t;
}

__global__ void k_iota(uint32_t* __restrict__ idx, uint32_t N) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < N) idx[p] = p;
}

// block-wide exclusive scan of one value per thread; returns the prefix and
// the block total.  `ws` holds one slot per wave plus one.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* ws, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) ws[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < BLOCK / 64; i++) { const uint32_t t = ws[i]; ws[i] = run; run += t; }
        ws[BLOCK / 64] = run;
    }
    __syncthreads();
    const uint32_t excl = ws[wid] + x - v;
    total = ws[BLOCK / 64];
    __syncthreads();
    return excl;
}

// ---------------------------------------------------------------------------
// one BVH level.  Per task: bin the primitives (3 axes x 64 bins; per-thread
// runs of equal bins are merged before the LDS atomics), evaluate SAH, choose
// leaf / SAH split / median split, partition stably.  Levels of a few big
// tasks (the top of the tree) spread each task over many blocks: partial bins
// per chunk, one block per task merges them and decides, and the partition
// runs chunk by chunk at offsets from the chunks' left counts.  Binning is a
// set of min / max / sum and the partition is stable, so both paths give the
// same tree.
// ---------------------------------------------------------------------------
// LDS state of one task's decision
struct LevelLds {
    uint32_t cnt[3][kBins];
    uint32_t b[3][kBins][12];   // lo.xyz, hi.xyz, clo.xyz, chi.xyz (ordered ints)
    uint32_t pcnt[3][kBins];    // primitives in bins 0..b
    float best_cost[3];
    int best_bin[3];
    float nb[6];                // node bounds
    int axis, bin, kind;
    uint32_t nleft;
};
static_assert(kBins == 64, "the SAH evaluation takes one bin per lane of a wave");
constexpr uint32_t kPartialWords = 3 * kBins * 13;   // counts, then bounds, of one chunk's bins

// a task's binning frame: centroid bounds low corner and bins per unit
struct TaskGeom {
    float cl[3], scale[3];
};
__device__ __forceinline__ TaskGeom task_geom(const BuildTask& task) {
    TaskGeom g;
    const float cl[3] = {task.clo.x, task.clo.y, task.clo.z};
    const float ch[3] = {task.chi.x, task.chi.y, task.chi.z};
    for (int k = 0; k < 3; k++) {
        const float ext = ch[k] - cl[k];
        g.cl[k] = cl[k];
        g.scale[k] = ext > 0.f ? (float)kBins * (1.f - 1e-6f) / ext : 0.f;
    }
    return g;
}

__device__ __forceinline__ bool bound_is_min(int k) { return k < 3 || (k >= 6 && k < 9); }

template <int BLOCK>
__device__ __forceinline__ void bins_clear(LevelLds& L) {
    for (int i = threadIdx.x; i < 3 * kBins; i += BLOCK) {
        const int ax = i / kBins, b = i % kBins;
        L.cnt[ax][b] = 0;
        for (int k = 0; k < 12; k++) L.b[ax][b][k] = bound_is_min(k) ? 0xFFFFFFFFu : 0u;
    }
}

// bins idx[r0, r1): each thread walks a contiguous run of the range
template <int BLOCK>
__device__ __forceinline__ void bins_add(LevelLds& L, const TaskGeom& g, const uint32_t* __restrict__ idx,
                                         const float4* __restrict__ plo, const float4* __restrict__ phi,
                                         uint32_t r0, uint32_t r1) {
    const uint32_t n = r1 - r0;
    const uint32_t chunk = (n + BLOCK - 1) / BLOCK;
    const uint32_t b0 = r0 + min(n, threadIdx.x * chunk);
    const uint32_t b1 = r0 + min(n, (threadIdx.x + 1) * chunk);
    int cur[3] = {-1, -1, -1};
    uint32_t cnt[3] = {0, 0, 0};
    uint32_t acc[3][12];
    auto flush = [&](int ax) {
        if (cur[ax] < 0 || cnt[ax] == 0) return;
        const int b = cur[ax];
        atomicAdd(&L.cnt[ax][b], cnt[ax]);
        for (int k = 0; k < 3; k++) {
            atomicMin(&L.b[ax][b][k], acc[ax][k]);
            atomicMax(&L.b[ax][b][3 + k], acc[ax][3 + k]);
            atomicMin(&L.b[ax][b][6 + k], acc[ax][6 + k]);
            atomicMax(&L.b[ax][b][9 + k], acc[ax][9 + k]);
        }
    };
    for (uint32_t i = b0; i < b1; i++) {
        const uint32_t p = idx[i];
        const float4 a = plo[p], h = phi[p];
        const float c[3] = {.5f * (a.x + h.x), .5f * (a.y + h.y), .5f * (a.z + h.z)};
        const uint32_t v[12] = {f2o(a.x), f2o(a.y), f2o(a.z), f2o(h.x), f2o(h.y), f2o(h.z),
                                f2o(c[0]), f2o(c[1]), f2o(c[2]), f2o(c[0]), f2o(c[1]), f2o(c[2])};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const int b = bin_of(c[ax], g.cl[ax], g.scale[ax]);
            if (b != cur[ax]) {
                flush(ax);
                cur[ax] = b;
                cnt[ax] = 0;
#pragma unroll
                for (int k = 0; k < 12; k++) acc[ax][k] = v[k];
            }
            cnt[ax]++;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                acc[ax][k] = min(acc[ax][k], v[k]);
                acc[ax][3 + k] = max(acc[ax][3 + k], v[3 + k]);
                acc[ax][6 + k] = min(acc[ax][6 + k], v[6 + k]);
                acc[ax][9 + k] = max(acc[ax][9 + k], v[9 + k]);
            }
        }
    }
    flush(0); flush(1); flush(2);
}

// SAH decision from the task's bins (every thread of the block calls it):
// the first wave evaluates the splits, one lane per bin -- inclusive prefix
// (bins 0..b) and suffix (bins b..63) bounds and counts by shuffles; min and
// max are exact, so they equal a bin-by-bin sweep.  The split after bin b
// costs area(prefix b) * count + area(suffix b+1) * count; the cheapest wins,
// the lowest bin on ties (the order of a sweep with a strict <).  Thread 0
// then writes the node's bounds and chooses leaf / SAH split / median split.
__device__ __forceinline__ void sah_decide(LevelLds& L, const BuildTask& task, uint32_t n,
                                           BNode* __restrict__ bnodes, uint32_t max_leaf) {
    if (threadIdx.x < 64) {
        const int b = (int)threadIdx.x;
        for (int ax = 0; ax < 3; ax++) {
            const uint32_t cb = L.cnt[ax][b];
            float pl[3], ph[3], sl[3], sh[3];
            for (int k = 0; k < 3; k++) {
                pl[k] = sl[k] = cb ? o2f(L.b[ax][b][k]) : FLT_MAX;
                ph[k] = sh[k] = cb ? o2f(L.b[ax][b][3 + k]) : -FLT_MAX;
            }
            uint32_t pc = cb, sc = cb;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const bool up = b >= off, dn = b + off < 64;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const float ul = __shfl_up(pl[k], off, 64), uh = __shfl_up(ph[k], off, 64);
                    const float dl = __shfl_down(sl[k], off, 64), dh = __shfl_down(sh[k], off, 64);
                    if (up) { pl[k] = fminf(pl[k], ul); ph[k] = fmaxf(ph[k], uh); }
                    if (dn) { sl[k] = fminf(sl[k], dl); sh[k] = fmaxf(sh[k], dh); }
                }
                const uint32_t uc = __shfl_up(pc, off, 64), dc = __shfl_down(sc, off, 64);
                if (up) pc += uc;
                if (dn) sc += dc;
            }
            float rl[3], rh[3];
#pragma unroll
            for (int k = 0; k < 3; k++) { rl[k] = __shfl_down(sl[k], 1, 64); rh[k] = __shfl_down(sh[k], 1, 64); }
            const uint32_t rc0 = __shfl_down(sc, 1, 64);
            const uint32_t rc = b < kBins - 1 ? rc0 : 0u;
            const float cost = area3(pl[0], pl[1], pl[2], ph[0], ph[1], ph[2]) * (float)pc +
                               area3(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]) * (float)rc;
            const bool valid = b < kBins - 1 && pc != 0 && rc != 0 && cost < FLT_MAX;
            // costs are >= 0, so their bits order like their values
            uint64_t key = valid ? ((uint64_t)__float_as_uint(cost) << 32) | (uint32_t)b : ~0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint64_t o = __shfl_xor(key, off, 64);
                key = o < key ? o : key;
            }
            L.pcnt[ax][b] = pc;
            if (b == 0) {
                L.best_cost[ax] = key == ~0ull ? FLT_MAX : __uint_as_float((uint32_t)(key >> 32));
                L.best_bin[ax] = key == ~0ull ? -1 : (int)(uint32_t)key;
            }
            if (ax == 0 && b == kBins - 1) {
                // node bounds: the union of the axis-0 bins
                for (int k = 0; k < 3; k++) { L.nb[k] = pl[k]; L.nb[3 + k] = ph[k]; }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float lx = L.nb[0], ly = L.nb[1], lz = L.nb[2], hx = L.nb[3], hy = L.nb[4], hz = L.nb[5];
        BNode nd;
        nd.lo = make_float4(lx, ly, lz, 0.f);
        nd.hi = make_float4(hx, hy, hz, 0.f);
        bnodes[task.node] = nd;
        int axis = -1;
        float best = FLT_MAX;
        for (int ax = 0; ax < 3; ax++)
            if (L.best_bin[ax] >= 0 && L.best_cost[ax] < best) { best = L.best_cost[ax]; axis = ax; }
        const float A = area3(lx, ly, lz, hx, hy, hz);
        const float leaf_cost = kCostIsect * (float)n * A;
        const float split_cost = kCostTrav * A + kCostIsect * best;
        int kind;   // 0 leaf, 1 SAH split, 2 median split
        if (n <= 1) kind = 0;
        else if (n <= max_leaf && (axis < 0 || leaf_cost <= split_cost)) kind = 0;
        else if (axis < 0 || task.depth >= (uint32_t)(kMaxTreeDepth - 28)) kind = 2;
        else kind = 1;
        L.kind = kind;
        L.axis = axis;
        L.bin = axis >= 0 ? L.best_bin[axis] : -1;
        L.nleft = kind == 1 ? L.pcnt[axis][L.bin] : n / 2;
    }
    __syncthreads();
}

// the task's split record: for a SAH split the centroid bounds of the two
// children (bins 0..bin and bin+1..63, reduced by the first wave); for a
// leaf or a median split by index the task's own bounds
__device__ __forceinline__ void split_out(const LevelLds& L, const BuildTask& task, SplitOut* __restrict__ out,
                                          uint32_t slot) {
    SplitOut so;
    so.kind = L.kind == 0 ? 0u : 1u;
    so.nleft = L.nleft;
    so.pad0 = so.pad1 = 0;
    if (L.kind == 1) {
        if (threadIdx.x < 64) {
            const int ax = L.axis, sb = L.bin, b = (int)threadIdx.x;
            const bool has = L.cnt[ax][b] != 0, left = b <= sb;
            float v[12];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float lo = o2f(L.b[ax][b][6 + k]), hi = o2f(L.b[ax][b][9 + k]);
                v[k] = has && left ? lo : FLT_MAX;
                v[3 + k] = has && left ? hi : -FLT_MAX;
                v[6 + k] = has && !left ? lo : FLT_MAX;
                v[9 + k] = has && !left ? hi : -FLT_MAX;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    v[k] = fminf(v[k], __shfl_xor(v[k], off, 64));
                    v[3 + k] = fmaxf(v[3 + k], __shfl_xor(v[3 + k], off, 64));
                    v[6 + k] = fminf(v[6 + k], __shfl_xor(v[6 + k], off, 64));
                    v[9 + k] = fmaxf(v[9 + k], __shfl_xor(v[9 + k], off, 64));
                }
            }
            if (b == 0) {
                so.clo_l = make_float4(v[0], v[1], v[2], 0.f);
                so.chi_l = make_float4(v[3], v[4], v[5], 0.f);
                so.clo_r = make_float4(v[6], v[7], v[8], 0.f);
                so.chi_r = make_float4(v[9], v[10], v[11], 0.f);
                out[slot] = so;
            }
        }
    } else if (threadIdx.x == 0) {
        so.clo_l = so.clo_r = task.clo;
        so.chi_l = so.chi_r = task.chi;
        out[slot] = so;
    }
}

// stable partition of idx[r0, r1) into tmp: the range's primitives left of
// the split go to tmp[lbase ...], the others to tmp[rbase ...], each in order
template <int BLOCK>
__device__ __forceinline__ void partition_range(const TaskGeom& g, int ax, int sb, const uint32_t* __restrict__ idx,
                                                uint32_t* __restrict__ tmp, const float4* __restrict__ plo,
                                                const float4* __restrict__ phi, uint32_t r0, uint32_t r1,
                                                uint32_t lbase, uint32_t rbase, uint32_t* ws) {
    const int PER = 4;
    const uint32_t n = r1 - r0;
    for (uint32_t t0 = 0; t0 < n; t0 += BLOCK * PER) {
        uint32_t vals[PER];
        uint32_t flags = 0, nl = 0;
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const uint32_t i = t0 + threadIdx.x * PER + e;
            vals[e] = 0;
            if (i < n) {
                const uint32_t p = idx[r0 + i];
                vals[e] = p;
                const float4 a = plo[p], h = phi[p];
                const float c = .5f * (comp4(a, ax) + comp4(h, ax));
                if (bin_of(c, g.cl[ax], g.scale[ax]) <= sb) { flags |= 1u << e; nl++; }
            }
        }
        uint32_t tot_l;
        const uint32_t pre_l = block_scan<BLOCK>(nl, ws, tot_l);
        const uint32_t cnt_here = (uint32_t)max(0, min(PER, (int)n - (int)(t0 + threadIdx.x * PER)));
        const uint32_t pre_all = min(n - t0, threadIdx.x * (uint32_t)PER);
        uint32_t l = lbase + pre_l, r = rbase + (pre_all - pre_l);
#pragma unroll
        for (int e = 0; e < PER; e++) {
            if ((uint32_t)e >= cnt_here) break;
            if (flags & (1u << e)) tmp[l++] = vals[e]; else tmp[r++] = vals[e];
        }
        const uint32_t tile_n = min((uint32_t)(BLOCK * PER), n - t0);
        lbase += tot_l;
        rbase += tile_n - tot_l;
    }
}

// one task on the whole block: bins, decision, stable partition, split record
template <int BLOCK>
__device__ __forceinline__ void level_task(LevelLds& L, uint32_t* s_ws, const BuildTask& task, uint32_t slot,
                                           const float4* __restrict__ plo, const float4* __restrict__ phi,
                                           uint32_t* __restrict__ idx, uint32_t* __restrict__ tmp,
                                           BNode* __restrict__ bnodes, SplitOut* __restrict__ out, uint32_t max_leaf) {
    const uint32_t n = task.end - task.begin;
    const TaskGeom g = task_geom(task);
    bins_clear<BLOCK>(L);
    __syncthreads();
    bins_add<BLOCK>(L, g, idx, plo, phi, task.begin, task.end);
    __syncthreads();
    sah_decide(L, task, n, bnodes, max_leaf);
    if (L.kind == 1) {
        partition_range<BLOCK>(g, L.axis, L.bin, idx, tmp, plo, phi, task.begin, task.end, task.begin,
                               task.begin + L.nleft, s_ws);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) idx[task.begin + i] = tmp[task.begin + i];
    }
    split_out(L, task, out, slot);
}

// one block per task
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_level(
        const BuildTask* __restrict__ tasks, const float4* __restrict__ plo,
        const float4* __restrict__ phi, uint32_t* __restrict__ idx, uint32_t* __restrict__ tmp,
        BNode* __restrict__ bnodes, SplitOut* __restrict__ out, uint32_t max_leaf) {
    __shared__ LevelLds L;
    __shared__ uint32_t s_ws[BLOCK / 64 + 1];
    level_task<BLOCK>(L, s_ws, tasks[blockIdx.x], blockIdx.x, plo, phi, idx, tmp, bnodes, out, max_leaf);
}

// ---------------------------------------------------------------------------
// Levels of small tasks: a thread per task of at most kTinyTask primitives,
// a wave per listed bigger task.  The thread makes the wave's decision: the
// same bins, prefix and suffix bounds as minima and maxima over the same
// primitives, the same cost expression and the same tie rule (the lowest
// bin).  It evaluates only the splits after an occupied bin: a split after
// an empty bin partitions like the split after the occupied bin below it and
// has the same cost at a higher bin, so it never wins.  (A minimum over
// zeros of both signs may keep the other sign than the wave's; the boxes and
// SAH costs are equal as numbers either way.)
#ifndef RTX_TINY_BITS
#define RTX_TINY_BITS 4
#endif
constexpr uint32_t kTinyBits = RTX_TINY_BITS;     // position bits of a sort key
constexpr uint32_t kTinyTask = 1u << kTinyBits;   // a power of two (the bitonic sort)
constexpr uint32_t kTinyLevel = 8192;  // tasks of a level below which a wave per task runs them all (fewer launches)
constexpr uint32_t kListGrid = 4096;   // waves of the listed-task kernel at most

__device__ __forceinline__ float centroid(const float4& a, const float4& h, int ax) {
    return .5f * (comp4(a, ax) + comp4(h, ax));
}

__device__ void tiny_task(const BuildTask& task, uint32_t slot, const float4* __restrict__ plo,
                          const float4* __restrict__ phi, uint32_t* __restrict__ idx, uint32_t* __restrict__ tmp,
                          BNode* __restrict__ bnodes, SplitOut* __restrict__ out, uint32_t max_leaf) {
    const uint32_t n = task.end - task.begin;
    const TaskGeom g = task_geom(task);
    const uint32_t* ix = idx + task.begin;
    float nlx = FLT_MAX, nly = FLT_MAX, nlz = FLT_MAX, nhx = -FLT_MAX, nhy = -FLT_MAX, nhz = -FLT_MAX;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = ix[i];
        const float4 a = plo[p], h = phi[p];
        nlx = fminf(nlx, a.x); nly = fminf(nly, a.y); nlz = fminf(nlz, a.z);
        nhx = fmaxf(nhx, h.x); nhy = fmaxf(nhy, h.y); nhz = fmaxf(nhz, h.z);
    }
    float best_cost[3];
    int best_bin[3];
    for (int ax = 0; ax < 3; ax++) {
        // the primitives in order of their bin (bin << kTinyBits | position, a bitonic
        // sort over kTinyTask register slots; absent slots sort last)
        uint32_t key[kTinyTask];
#pragma unroll
        for (uint32_t i = 0; i < kTinyTask; i++) {
            key[i] = 0xFFFFFFFFu;
            if (i < n) {
                const uint32_t p = ix[i];
                key[i] = ((uint32_t)bin_of(centroid(plo[p], phi[p], ax), g.cl[ax], g.scale[ax]) << kTinyBits) | i;
            }
        }
#pragma unroll
        for (uint32_t k = 2; k <= kTinyTask; k <<= 1) {
#pragma unroll
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (uint32_t i = 0; i < kTinyTask; i++) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const uint32_t a = key[i], b = key[l];
                        const bool lo_first = (i & k) == 0;
                        key[i] = lo_first ? min(a, b) : max(a, b);
                        key[l] = lo_first ? max(a, b) : min(a, b);
                    }
                }
            }
        }
        // prefix terms area(bins <= bin of position k) * count and suffix terms
        // area(bins of positions >= k) * count, as the wave's scans form them
        float tl[kTinyTask], tr[kTinyTask];
        float lx0 = FLT_MAX, ly0 = FLT_MAX, lz0 = FLT_MAX, lx1 = -FLT_MAX, ly1 = -FLT_MAX, lz1 = -FLT_MAX;
#pragma unroll
        for (uint32_t k = 0; k < kTinyTask; k++) {
            tl[k] = 0.f;
            if (k < n) {
                const uint32_t p = ix[key[k] & (kTinyTask - 1u)];
                const float4 a = plo[p], h = phi[p];
                lx0 = fminf(lx0, a.x); ly0 = fminf(ly0, a.y); lz0 = fminf(lz0, a.z);
                lx1 = fmaxf(lx1, h.x); ly1 = fmaxf(ly1, h.y); lz1 = fmaxf(lz1, h.z);
                tl[k] = area3(lx0, ly0, lz0, lx1, ly1, lz1) * (float)(k + 1);
            }
        }
        float rx0 = FLT_MAX, ry0 = FLT_MAX, rz0 = FLT_MAX, rx1 = -FLT_MAX, ry1 = -FLT_MAX, rz1 = -FLT_MAX;
#pragma unroll
        for (int k = (int)kTinyTask - 1; k >= 0; k--) {
            tr[k] = 0.f;
            if ((uint32_t)k < n) {
                const uint32_t p = ix[key[k] & (kTinyTask - 1u)];
                const float4 a = plo[p], h = phi[p];
                rx0 = fminf(rx0, a.x); ry0 = fminf(ry0, a.y); rz0 = fminf(rz0, a.z);
                rx1 = fmaxf(rx1, h.x); ry1 = fmaxf(ry1, h.y); rz1 = fmaxf(rz1, h.z);
                tr[k] = area3(rx0, ry0, rz0, rx1, ry1, rz1) * (float)(n - (uint32_t)k);
            }
        }
        // a split after bin b of position k: position k + 1 holds a higher bin
        uint64_t best = ~0ull;
#pragma unroll
        for (uint32_t k = 0; k + 1 < kTinyTask; k++) {
            if (k + 1 < n && (key[k] >> kTinyBits) != (key[k + 1] >> kTinyBits)) {
                const float cost = tl[k] + tr[k + 1];
                if (cost < FLT_MAX) {
                    const uint64_t c = ((uint64_t)__float_as_uint(cost) << 32) | (key[k] >> kTinyBits);
                    best = c < best ? c : best;
                }
            }
        }
        best_cost[ax] = best == ~0ull ? FLT_MAX : __uint_as_float((uint32_t)(best >> 32));
        best_bin[ax] = best == ~0ull ? -1 : (int)(uint32_t)best;
    }
    BNode nd;
    nd.lo = make_float4(nlx, nly, nlz, 0.f);
    nd.hi = make_float4(nhx, nhy, nhz, 0.f);
    bnodes[task.node] = nd;
    int axis = -1;
    float best = FLT_MAX;
    for (int ax = 0; ax < 3; ax++)
        if (best_bin[ax] >= 0 && best_cost[ax] < best) { best = best_cost[ax]; axis = ax; }
    const float A = area3(nlx, nly, nlz, nhx, nhy, nhz);
    const float leaf_cost = kCostIsect * (float)n * A;
    const float split_cost = kCostTrav * A + kCostIsect * best;
    int kind;   // 0 leaf, 1 SAH split, 2 median split (as in sah_decide)
    if (n <= 1) kind = 0;
    else if (n <= max_leaf && (axis < 0 || leaf_cost <= split_cost)) kind = 0;
    else if (axis < 0 || task.depth >= (uint32_t)(kMaxTreeDepth - 28)) kind = 2;
    else kind = 1;
    SplitOut so;
    so.kind = kind == 0 ? 0u : 1u;
    so.pad0 = so.pad1 = 0;
    if (kind == 1) {
        // stable partition into tmp and back, with the centroid bounds of both sides
        const int sb = best_bin[axis];
        float llx = FLT_MAX, lly = FLT_MAX, llz = FLT_MAX, lhx = -FLT_MAX, lhy = -FLT_MAX, lhz = -FLT_MAX;
        float rlx = FLT_MAX, rly = FLT_MAX, rlz = FLT_MAX, rhx = -FLT_MAX, rhy = -FLT_MAX, rhz = -FLT_MAX;
        uint32_t nl = 0;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t p = ix[i];
            const float4 a = plo[p], h = phi[p];
            const float cx = centroid(a, h, 0), cy = centroid(a, h, 1), cz = centroid(a, h, 2);
            const float ca = axis == 0 ? cx : (axis == 1 ? cy : cz);
            if (bin_of(ca, g.cl[axis], g.scale[axis]) <= sb) {
                llx = fminf(llx, cx); lly = fminf(lly, cy); llz = fminf(llz, cz);
                lhx = fmaxf(lhx, cx); lhy = fmaxf(lhy, cy); lhz = fmaxf(lhz, cz);
                tmp[task.begin + nl++] = p;
            } else {
                rlx = fminf(rlx, cx); rly = fminf(rly, cy); rlz = fminf(rlz, cz);
                rhx = fmaxf(rhx, cx); rhy = fmaxf(rhy, cy); rhz = fmaxf(rhz, cz);
            }
        }
        uint32_t r = nl;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t p = ix[i];
            if (bin_of(centroid(plo[p], phi[p], axis), g.cl[axis], g.scale[axis]) > sb) tmp[task.begin + r++] = p;
        }
        for (uint32_t i = 0; i < n; i++) idx[task.begin + i] = tmp[task.begin + i];
        so.nleft = nl;
        so.clo_l = make_float4(llx, lly, llz, 0.f);
        so.chi_l = make_float4(lhx, lhy, lhz, 0.f);
        so.clo_r = make_float4(rlx, rly, rlz, 0.f);
        so.chi_r = make_float4(rhx, rhy, rhz, 0.f);
    } else {
        so.nleft = n / 2;
        so.clo_l = so.clo_r = task.clo;
        so.chi_l = so.chi_r = task.chi;
    }
    out[slot] = so;
}

// a thread per task; tasks above kTinyTask primitives are appended to `big`
// (one atomic per wave: their order in the list does not matter, each task
// writes only its own slot and primitive range)
__global__ __launch_bounds__(256) void k_level_tiny(const BuildTask* __restrict__ tasks, uint32_t ntasks,
                                                    const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                    uint32_t* __restrict__ idx, uint32_t* __restrict__ tmp,
                                                    BNode* __restrict__ bnodes, SplitOut* __restrict__ out,
                                                    uint32_t max_leaf, uint32_t* __restrict__ big,
                                                    uint32_t* __restrict__ nbig) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool is_big = false;
    if (t < ntasks) {
        const BuildTask task = tasks[t];
        if (task.end - task.begin > kTinyTask) is_big = true;
        else tiny_task(task, t, plo, phi, idx, tmp, bnodes, out, max_leaf);
    }
    const uint64_t m = __ballot(is_big);
    if (m) {
        const uint32_t lane = threadIdx.x & 63u;
        const int first = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane == (uint32_t)first) base = atomicAdd(nbig, (uint32_t)__popcll(m));
        base = __shfl(base, first, 64);
        if (is_big) big[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = t;
    }
}

// a wave per listed task, grid-strided over the list
__global__ __launch_bounds__(64) void k_level_list(const BuildTask* __restrict__ tasks,
                                                   const uint32_t* __restrict__ big, const uint32_t* __restrict__ nbig,
                                                   const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                   uint32_t* __restrict__ idx, uint32_t* __restrict__ tmp,
                                                   BNode* __restrict__ bnodes, SplitOut* __restrict__ out,
                                                   uint32_t max_leaf) {
    __shared__ LevelLds L;
    __shared__ uint32_t s_ws[2];
    const uint32_t nb = *nbig;
    for (uint32_t k = blockIdx.x; k < nb; k += gridDim.x) {
        const uint32_t t = big[k];
        level_task<64>(L, s_ws, tasks[t], t, plo, phi, idx, tmp, bnodes, out, max_leaf);
        __syncthreads();
    }
}

// big tasks, step 1: block (chunk c, task t) bins chunk c of the task into
// partial[t * G + c]
__device__ __forceinline__ void chunk_range(const BuildTask& task, uint32_t G, uint32_t c, uint32_t& r0,
                                            uint32_t& r1) {
    const uint32_t n = task.end - task.begin, chunk = (n + G - 1) / G;
    r0 = task.begin + min(n, c * chunk);
    r1 = task.begin + min(n, (c + 1) * chunk);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bin_wide(const BuildTask* __restrict__ tasks, const float4* __restrict__ plo,
                                                    const float4* __restrict__ phi, const uint32_t* __restrict__ idx,
                                                    uint32_t G, uint32_t* __restrict__ partial) {
    __shared__ LevelLds L;
    const BuildTask task = tasks[blockIdx.y];
    uint32_t r0, r1;
    chunk_range(task, G, blockIdx.x, r0, r1);
    bins_clear<BLOCK>(L);
    __syncthreads();
    bins_add<BLOCK>(L, task_geom(task), idx, plo, phi, r0, r1);
    __syncthreads();
    uint32_t* P = partial + (size_t)(blockIdx.y * G + blockIdx.x) * kPartialWords;
    const uint32_t* c = &L.cnt[0][0];
    const uint32_t* bb = &L.b[0][0][0];
    for (uint32_t i = threadIdx.x; i < 3 * kBins; i += BLOCK) P[i] = c[i];
    for (uint32_t i = threadIdx.x; i < 3 * kBins * 12; i += BLOCK) P[3 * kBins + i] = bb[i];
}

// big tasks, step 2a: the chunks' bins of each task merged into one set
// (sums of counts, minima and maxima of bounds: exact, in any order).
// Block (x, t) merges words 16x .. 16x+15 of task t: 16 groups of 16
// threads each reduce every 16th chunk, then the first group the groups.
constexpr uint32_t kMergeWords = 16;
static_assert(kPartialWords % kMergeWords == 0, "whole blocks of merged words");
__global__ __launch_bounds__(256) void k_merge_wide(uint32_t G, const uint32_t* __restrict__ partial,
                                                    uint32_t* __restrict__ merged) {
    __shared__ uint32_t s_v[256 / kMergeWords][kMergeWords + 1];
    const uint32_t w = threadIdx.x % kMergeWords, grp = threadIdx.x / kMergeWords;
    const uint32_t i = blockIdx.x * kMergeWords + w;
    // word kind: 0 count (sum), 1 lower bound (min), 2 upper bound (max)
    const int kind = i < 3 * kBins ? 0 : (bound_is_min((int)((i - 3 * kBins) % 12)) ? 1 : 2);
    const uint32_t* P = partial + (size_t)blockIdx.y * G * kPartialWords + i;
    uint32_t v = kind == 1 ? 0xFFFFFFFFu : 0u;
    for (uint32_t k = grp; k < G; k += 256 / kMergeWords) {
        const uint32_t x = P[(size_t)k * kPartialWords];
        v = kind == 0 ? v + x : (kind == 1 ? min(v, x) : max(v, x));
    }
    s_v[grp][w] = v;
    __syncthreads();
    if (grp == 0) {
        for (uint32_t g = 1; g < 256 / kMergeWords; g++) {
            const uint32_t x = s_v[g][w];
            v = kind == 0 ? v + x : (kind == 1 ? min(v, x) : max(v, x));
        }
        merged[(size_t)blockIdx.y * kPartialWords + i] = v;
    }
}

// big tasks, step 2b: one block per task reads its G sets of bins (one, after
// k_merge_wide), merges them and decides as k_level does; wsplit[t] = (kind,
// axis, bin, nleft) for the partition
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_decide_wide(const BuildTask* __restrict__ tasks, uint32_t G,
                                                       const uint32_t* __restrict__ partial,
                                                       BNode* __restrict__ bnodes, SplitOut* __restrict__ out,
                                                       uint4* __restrict__ wsplit, uint32_t max_leaf) {
    __shared__ LevelLds L;
    const BuildTask task = tasks[blockIdx.x];
    const uint32_t* P = partial + (size_t)blockIdx.x * G * kPartialWords;
    uint32_t* c = &L.cnt[0][0];
    uint32_t* bb = &L.b[0][0][0];
    for (uint32_t i = threadIdx.x; i < 3 * kBins; i += BLOCK) {
        uint32_t s = 0;
        for (uint32_t k = 0; k < G; k++) s += P[(size_t)k * kPartialWords + i];
        c[i] = s;
    }
    for (uint32_t i = threadIdx.x; i < 3 * kBins * 12; i += BLOCK) {
        const bool mn = bound_is_min((int)(i % 12));
        uint32_t v = mn ? 0xFFFFFFFFu : 0u;
        for (uint32_t k = 0; k < G; k++) {
            const uint32_t x = P[(size_t)k * kPartialWords + 3 * kBins + i];
            v = mn ? min(v, x) : max(v, x);
        }
        bb[i] = v;
    }
    __syncthreads();
    sah_decide(L, task, task.end - task.begin, bnodes, max_leaf);
    split_out(L, task, out, blockIdx.x);
    if (threadIdx.x == 0) wsplit[blockIdx.x] = make_uint4((uint32_t)L.kind, (uint32_t)L.axis, (uint32_t)L.bin, L.nleft);
}

// big tasks, step 3: primitives left of the split in each chunk
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_count_wide(const BuildTask* __restrict__ tasks, const float4* __restrict__ plo,
                                                      const float4* __restrict__ phi, const uint32_t* __restrict__ idx,
                                                      uint32_t G, const uint4* __restrict__ wsplit,
                                                      uint32_t* __restrict__ lcount) {
    __shared__ uint32_t s_n;
    const uint4 w = wsplit[blockIdx.y];
    if (w.x != 1u) return;   // the whole block: no SAH split
    const BuildTask task = tasks[blockIdx.y];
    const TaskGeom g = task_geom(task);
    uint32_t r0, r1;
    chunk_range(task, G, blockIdx.x, r0, r1);
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int ax = (int)w.y, sb = (int)w.z;
    uint32_t nl = 0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        const uint32_t p = idx[i];
        const float4 a = plo[p], h = phi[p];
        if (bin_of(.5f * (comp4(a, ax) + comp4(h, ax)), g.cl[ax], g.scale[ax]) <= sb) nl++;
    }
    atomicAdd(&s_n, nl);
    __syncthreads();
    if (threadIdx.x == 0) lcount[blockIdx.y * G + blockIdx.x] = s_n;
}

// big tasks, step 4: each chunk partitions stably into tmp after the left
// and right primitives of the chunks before it
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_scatter_wide(const BuildTask* __restrict__ tasks, const float4* __restrict__ plo,
                                                        const float4* __restrict__ phi, const uint32_t* __restrict__ idx,
                                                        uint32_t* __restrict__ tmp, uint32_t G,
                                                        const uint4* __restrict__ wsplit,
                                                        const uint32_t* __restrict__ lcount) {
    __shared__ uint32_t s_ws[BLOCK / 64 + 1];
    __shared__ uint32_t s_lb;
    const uint4 w = wsplit[blockIdx.y];
    if (w.x != 1u) return;
    const BuildTask task = tasks[blockIdx.y];
    uint32_t r0, r1;
    chunk_range(task, G, blockIdx.x, r0, r1);
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t k = 0; k < blockIdx.x; k++) s += lcount[blockIdx.y * G + k];
        s_lb = s;
    }
    __syncthreads();
    const uint32_t lb = s_lb, before = r0 - task.begin;
    partition_range<BLOCK>(task_geom(task), (int)w.y, (int)w.z, idx, tmp, plo, phi, r0, r1, task.begin + lb,
                           task.begin + w.w + (before - lb), s_ws);
}

// big tasks, step 5: the partitioned ranges back into idx
__global__ void k_copy_wide(const BuildTask* __restrict__ tasks, uint32_t G, const uint4* __restrict__ wsplit,
                            const uint32_t* __restrict__ tmp, uint32_t* __restrict__ idx) {
    if (wsplit[blockIdx.y].x != 1u) return;
    uint32_t r0, r1;
    chunk_range(tasks[blockIdx.y], G, blockIdx.x, r0, r1);
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) idx[i] = tmp[i];
}

// single-block exclusive scan of `n` uints (kind flags) -> out, total -> *total
__global__ __launch_bounds__(kScanBlock) void k_scan_kind(const SplitOut* __restrict__ so, uint32_t n,
                                                         uint32_t* __restrict__ out, uint32_t* __restrict__ total) {
    __shared__ uint32_t s_ws[kScanBlock / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? so[i].kind : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan<kScanBlock>(v, s_ws, tot);
        if (i < n) out[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// the same scan over many blocks, for long arrays: tile sums, a scan of
// them (k_scan_u32), then each tile rescanned from its offset
constexpr uint32_t kScanTile = 4 * kScanBlock;
// the values to scan: the split flags of a level, or plain counts
struct KindOf {
    const SplitOut* so;
    __device__ __forceinline__ uint32_t operator()(uint32_t i) const { return so[i].kind; }
};
struct U32Of {
    const uint32_t* p;
    __device__ __forceinline__ uint32_t operator()(uint32_t i) const { return p[i]; }
};
template <class Get>
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(Get get, uint32_t n, uint32_t* __restrict__ tsum) {
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    uint32_t c = 0;
    const uint32_t b0 = blockIdx.x * kScanTile;
    for (uint32_t i = b0 + threadIdx.x; i < n && i < b0 + kScanTile; i += kScanBlock) c += get(i);
    atomicAdd(&s_n, c);
    __syncthreads();
    if (threadIdx.x == 0) tsum[blockIdx.x] = s_n;
}
template <class Get>
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(Get get, uint32_t n, const uint32_t* __restrict__ toff,
                                                          uint32_t* __restrict__ out) {
    __shared__ uint32_t s_ws[kScanBlock / 64 + 1];
    uint32_t carry = toff[blockIdx.x];
    const uint32_t b0 = blockIdx.x * kScanTile;
    for (uint32_t base = b0; base < b0 + kScanTile && base < n; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? get(i) : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan<kScanBlock>(v, s_ws, tot);
        if (i < n) out[i] = carry + ex;
        carry += tot;
    }
}

__global__ void k_emit(const BuildTask* __restrict__ tasks, const SplitOut* __restrict__ so,
                       const uint32_t* __restrict__ rank, uint32_t ntasks, uint32_t node_base,
                       BNode* __restrict__ bnodes, BuildTask* __restrict__ next,
                       BuildStats* __restrict__ st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t leaf_depth = 0, leaf_n = 0;   // of a leaf, for the statistics
    bool leaf = false;
    if (i < ntasks) {
        const BuildTask t = tasks[i];
        const SplitOut s = so[i];
        const uint32_t n = t.end - t.begin;
        if (s.kind == 0) {
            bnodes[t.node].lo.w = __uint_as_float(t.begin);
            bnodes[t.node].hi.w = __uint_as_float(n);
            leaf = true;
            leaf_depth = t.depth + 1;
            leaf_n = n;
        } else {
            const uint32_t L = node_base + 2 * rank[i];
            bnodes[t.node].lo.w = __uint_as_float(L);
            bnodes[t.node].hi.w = __uint_as_float(0u);
            BuildTask a, b;
            a.node = L; a.begin = t.begin; a.end = t.begin + s.nleft; a.depth = t.depth + 1;
            a.clo = s.clo_l; a.chi = s.chi_l;
            b.node = L + 1; b.begin = t.begin + s.nleft; b.end = t.end; b.depth = t.depth + 1;
            b.clo = s.clo_r; b.chi = s.chi_r;
            next[2 * rank[i]] = a;
            next[2 * rank[i] + 1] = b;
        }
    }
    // leaf count, deepest leaf and largest leaf with one atomic each per wave
    // (millions of leaves on one address otherwise serialise in L2)
    const uint64_t m = __ballot(leaf);
    if (m == 0) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        leaf_depth = max(leaf_depth, (uint32_t)__shfl_xor(leaf_depth, off, 64));
        leaf_n = max(leaf_n, (uint32_t)__shfl_xor(leaf_n, off, 64));
    }
    if ((threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)m) - 1)) {
        atomicAdd(&st->leaves, (uint32_t)__popcll(m));
        atomicMax(&st->max_depth, leaf_depth);
        atomicMax(&st->max_leaf, leaf_n);
    }
}

// SAH cost of the finished tree (relative to the root area)
// SAH cost of the binary tree (statistics): grid-strided partial sums, a wave
// reduction, then one atomic per block (the order of the double sums varies
// from run to run in the last bits)
constexpr uint32_t kSahBlocks = 1024;
__global__ __launch_bounds__(256) void k_sah(const BNode* __restrict__ bnodes, uint32_t nnodes,
                                             BuildStats* __restrict__ st) {
    __shared__ double s_v[256 / 64];
    double v = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nnodes; i += gridDim.x * blockDim.x) {
        const BNode nd = bnodes[i];
        const float a = area3(nd.lo.x, nd.lo.y, nd.lo.z, nd.hi.x, nd.hi.y, nd.hi.z);
        const uint32_t cnt = __float_as_uint(nd.hi.w);
        v += cnt ? (double)kCostIsect * cnt * a : (double)kCostTrav * a;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_v[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 256 / 64; w++) t += s_v[w];
        if (t != 0.0) atomicAdd(&st->sah, t);
    }
}

__global__ void k_inner_flag(const BNode* __restrict__ bnodes, uint32_t nnodes, SplitOut* __restrict__ flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnodes) flag[i].kind = __float_as_uint(bnodes[i].hi.w) == 0u ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// collapse of the binary tree into 4-wide nodes, level by level from the root
// (the grouping of builds with RTX_DP_COLLAPSE=0; the default is the
// SAH-optimal one below).  A wide node starts from a binary node's two
// children and repeatedly opens the inner child of largest surface area until
// it has 4 children or only leaves; its inner children become the wide nodes
// of the next level.
__global__ void k_collapse4(const BNode* __restrict__ bn, const uint32_t* __restrict__ front, uint32_t n,
                            uint4* __restrict__ cid, uint32_t* __restrict__ ninner) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t L = __float_as_uint(bn[front[i]].lo.w);
    uint32_t c0 = L, c1 = L + 1, c2 = kNoChild, c3 = kNoChild;
    uint32_t cnt = 2;
    for (int it = 0; it < 2; it++) {
        // inner child of largest area among c0..c(cnt-1)
        float best = -1.f;
        int bk = -1;
        for (int k = 0; k < 4; k++) {
            const uint32_t c = k == 0 ? c0 : k == 1 ? c1 : k == 2 ? c2 : c3;
            if ((uint32_t)k >= cnt) break;
            const BNode b = bn[c];
            if (__float_as_uint(b.hi.w) != 0u) continue;    // leaf
            const float a = area3(b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z);
            if (a > best) { best = a; bk = k; }
        }
        if (bk < 0) break;
        const uint32_t opened = bk == 0 ? c0 : bk == 1 ? c1 : bk == 2 ? c2 : c3;
        const uint32_t Lb = __float_as_uint(bn[opened].lo.w);
        if (bk == 0) c0 = Lb; else if (bk == 1) c1 = Lb; else if (bk == 2) c2 = Lb; else c3 = Lb;
        if (cnt == 2) c2 = Lb + 1; else c3 = Lb + 1;
        cnt++;
    }
    uint32_t m = 0;
    for (int k = 0; k < 4; k++) {
        const uint32_t c = k == 0 ? c0 : k == 1 ? c1 : k == 2 ? c2 : c3;
        if (c != kNoChild && __float_as_uint(bn[c].hi.w) == 0u) m++;
    }
    cid[i] = make_uint4(c0, c1, c2, c3);
    ninner[i] = m;
}

// ---------------------------------------------------------------------------
// SAH-optimal collapse (after Ylitie, Karras and Laine 2017, "Efficient
// incoherent ray traversal on GPUs through compressed wide BVHs", for 4-wide
// nodes).  D(n, k) is the least SAH cost of the binary subtree at n as a
// forest of at most k roots, each root a leaf (all the subtree's primitives,
// at most max_leaf, contiguous in the leaf order) or a 4-wide node whose
// children are a forest of at most 4 roots under it:
//   D(n, 1) = min(A(n) * n_prims * C_isect  [n_prims <= max_leaf],
//                 A(n) * C_trav + min_i D(L, i) + D(R, 4 - i))
//   D(n, k) = min(D(n, 1), min_{0<i<k} D(L, i) + D(R, k - i))
// bottom-up, one binary level at a time (the build allocates the nodes level
// by level, so a level's children are all in the next one).  A binary leaf
// stays a leaf; the root stays a 4-wide node.  The collapse then takes each
// wide node's children from the choices (split of the 4 roots between L and
// R, then each subtree's own split).  The binary tree and its statistics are
// unchanged; only the grouping into 4-wide nodes and leaves is.
//
// meta bits: 0 the subtree is one leaf; 1-2 the split of 4 roots between L
// and R (1..3 to L); 3-4, 5-6, 7-8 the split for k = 2, 3, 4 (0: the subtree
// itself is the one root, else the roots given to L)
__global__ void k_dp_level(const BNode* __restrict__ bn, uint32_t begin, uint32_t n, float4* __restrict__ dc,
                           uint32_t* __restrict__ ntri, uint32_t* __restrict__ first, uint32_t* __restrict__ meta,
                           uint32_t max_leaf) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = begin + t;
    const BNode b = bn[i];
    const float A = area3(b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z);
    const uint32_t cnt = __float_as_uint(b.hi.w);
    if (cnt) {
        const float c = A * (float)cnt * kCostIsect;
        dc[i] = make_float4(c, c, c, c);
        ntri[i] = cnt;
        first[i] = __float_as_uint(b.lo.w);
        meta[i] = 1u;
        return;
    }
    const uint32_t L = __float_as_uint(b.lo.w), R = L + 1u;
    const float4 dl = dc[L], dr = dc[R];
    const float l[5] = {0.f, dl.x, dl.y, dl.z, dl.w}, r[5] = {0.f, dr.x, dr.y, dr.z, dr.w};
    float d2[5];
    uint32_t b2[5];
    for (int k = 2; k <= 4; k++) {
        float best = l[1] + r[k - 1];
        uint32_t bi = 1;
        for (int s = 2; s < k; s++) {
            const float c = l[s] + r[k - s];
            if (c < best) { best = c; bi = (uint32_t)s; }
        }
        d2[k] = best;
        b2[k] = bi;
    }
    const uint32_t nt = ntri[L] + ntri[R];
    const float cnode = A * kCostTrav + d2[4];
    const bool leaf = i != 0u && nt <= max_leaf && A * (float)nt * kCostIsect <= cnode;
    const float d1 = leaf ? A * (float)nt * kCostIsect : cnode;
    uint32_t m = (leaf ? 1u : 0u) | (b2[4] << 1);
    float d[5];
    d[1] = d1;
    for (int k = 2; k <= 4; k++) {
        const bool self = d1 <= d2[k];
        d[k] = self ? d1 : d2[k];
        m |= (self ? 0u : b2[k]) << (3 + 2 * (k - 2));
    }
    dc[i] = make_float4(d[1], d[2], d[3], d[4]);
    ntri[i] = nt;
    first[i] = first[L];
    meta[i] = m;
}

// the children of the wide node at binary node front[i]: the forest of 4
// roots that the choices give under it (left subtree first)
__global__ void k_collapse4_dp(const BNode* __restrict__ bn, const uint32_t* __restrict__ front, uint32_t n,
                               const uint32_t* __restrict__ meta, uint4* __restrict__ cid,
                               uint32_t* __restrict__ ninner) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = front[i];
    const uint32_t L = __float_as_uint(bn[b].lo.w);
    const uint32_t kl = (meta[b] >> 1) & 3u;
    // (subtree, roots) pairs still to expand; their roots add up to 4, so
    // neither the stack nor the output holds more than 4
    uint32_t sn[4], sk[4], out[4] = {kNoChild, kNoChild, kNoChild, kNoChild};
    int sp = 0, no = 0;
    sn[sp] = L + 1u; sk[sp++] = 4u - kl;
    sn[sp] = L; sk[sp++] = kl;
    while (sp > 0) {
        --sp;
        const uint32_t x = sn[sp], k = sk[sp];
        const uint32_t mx = meta[x];
        const uint32_t s = k == 1u ? 0u : (mx >> (3u + 2u * (k - 2u))) & 3u;
        if (s == 0u) {
            out[no++] = x;
        } else {
            const uint32_t XL = __float_as_uint(bn[x].lo.w);
            sn[sp] = XL + 1u; sk[sp++] = k - s;
            sn[sp] = XL; sk[sp++] = s;
        }
    }
    uint32_t m = 0;
    for (int k = 0; k < no; k++) m += (meta[out[k]] & 1u) ? 0u : 1u;
    cid[i] = make_uint4(out[0], out[1], out[2], out[3]);
    ninner[i] = m;
}

// Writes the wide nodes of one level.  With meta (the SAH-optimal collapse)
// a child is a leaf when its subtree is one (primitives first[c] ..
// first[c] + ntri[c] - 1 of the leaf order); without, when it is a binary leaf.
__global__ void k_write4(const BNode* __restrict__ bn, const uint4* __restrict__ cid, const uint32_t* __restrict__ off,
                         uint32_t n, uint32_t base_this, uint32_t base_next, uint32_t empty_ref,
                         float4* __restrict__ nodes4, uint32_t* __restrict__ next_front,
                         const uint32_t* __restrict__ meta, const uint32_t* __restrict__ ntri,
                         const uint32_t* __restrict__ first) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 c4 = cid[i];
    float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
    uint32_t ref[4];
    uint32_t j = off[i];
    for (int k = 0; k < 4; k++) {
        const uint32_t c = k == 0 ? c4.x : k == 1 ? c4.y : k == 2 ? c4.z : c4.w;
        if (c == kNoChild) {
            // an empty slot: a point box far outside any scene, which a slab
            // test accepts only for rays aimed at it with an unbounded tmax;
            // its ref is a one-triangle leaf holding a degenerate triangle
            // (no ray hits it), so even then the traversal stays in bounds
            lx[k] = hx[k] = ly[k] = hy[k] = lz[k] = hz[k] = kFarAway;
            ref[k] = empty_ref;
            continue;
        }
        const BNode b = bn[c];
        lx[k] = b.lo.x; hx[k] = b.hi.x; ly[k] = b.lo.y; hy[k] = b.hi.y; lz[k] = b.lo.z; hz[k] = b.hi.z;
        const uint32_t cntp = __float_as_uint(b.hi.w);
        if (meta && (meta[c] & 1u)) {
            ref[k] = leaf_ref(first[c], ntri[c]);
        } else if (!meta && cntp) {
            ref[k] = leaf_ref(__float_as_uint(b.lo.w), cntp);
        } else {
            ref[k] = base_next + j;
            next_front[j] = c;
            j++;
        }
    }
    float4* o = nodes4 + kWideNodeF4 * (size_t)(base_this + i);
    o[0] = make_float4(lx[0], lx[1], lx[2], lx[3]);
    o[1] = make_float4(hx[0], hx[1], hx[2], hx[3]);
    o[2] = make_float4(ly[0], ly[1], ly[2], ly[3]);
    o[3] = make_float4(hy[0], hy[1], hy[2], hy[3]);
    o[4] = make_float4(lz[0], lz[1], lz[2], lz[3]);
    o[5] = make_float4(hz[0], hz[1], hz[2], hz[3]);
    o[6] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]),
                       __uint_as_float(ref[3]));
    o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// single-block exclusive scan of n uints -> out, total -> *total
__global__ __launch_bounds__(kScanBlock) void k_scan_u32(const uint32_t* __restrict__ in, uint32_t n,
                                                        uint32_t* __restrict__ out, uint32_t* __restrict__ total) {
    __shared__ uint32_t s_ws[kScanBlock / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? in[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan<kScanBlock>(v, s_ws, tot);
        if (i < n) out[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_gather_tris(const float4* __restrict__ tri_world, const uint32_t* __restrict__ idx, uint32_t N,
                              float4* __restrict__ tris) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == N) {
        // the degenerate triangle behind the empty slots of 4-wide nodes
        const float4 p = make_float4(kFarAway, kFarAway, kFarAway, __uint_as_float(kNoHit));
        tris[3 * (size_t)N + 0] = p;
        tris[3 * (size_t)N + 1] = make_float4(kFarAway, kFarAway, kFarAway, 0.f);
        tris[3 * (size_t)N + 2] = make_float4(kFarAway, kFarAway, kFarAway, __uint_as_float(kNoHit));
        return;
    }
    if (k > N) return;
    const size_t p = idx[k];
    tris[3 * (size_t)k + 0] = tri_world[3 * p + 0];
    tris[3 * (size_t)k + 1] = tri_world[3 * p + 1];
    tris[3 * (size_t)k + 2] = tri_world[3 * p + 2];
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    T* release() { T* q = p; p = nullptr; return q; }
};

#define RTX_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) { if (err) *err = std::string(#expr) + ": " + hipGetErrorString(e_); return e_; } \
    } while (0)

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

hipError_t flatten_scene(const InstBuild* d_insts, uint32_t ninst, const MeshRef* d_meshes, uint32_t N,
                         float4* d_tri_world, float4* d_plo, float4* d_phi, hipStream_t s) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flatten, dim3(blocks_for(N, 256)), dim3(256), 0, s, d_insts, ninst, d_meshes, N,
                       d_tri_world, d_plo, d_phi);
    return hipGetLastError();
}

hipError_t build_bvh(const float4* d_tri_world, const float4* d_plo, const float4* d_phi, uint32_t N,
                     hipStream_t s, BuiltBVH* outp, std::string* err, const BvhOptions& opt) {
    BuiltBVH out;
    const uint32_t max_leaf = opt.max_leaf < 1 ? 1u : (opt.max_leaf > kMaxLeaf ? kMaxLeaf : opt.max_leaf);
    if (N == 0) {
        out.root = kEmptyRoot;
        *outp = out;
        return hipSuccess;
    }
    DevBuf<uint32_t> idx, tmp, rank, total;
    DevBuf<BuildTask> tasks_a, tasks_b;
    DevBuf<SplitOut> so;
    DevBuf<BNode> bnodes;
    DevBuf<BuildStats> st;
    const size_t max_nodes = 2 * (size_t)N;
    RTX_TRY(idx.alloc(N));
    RTX_TRY(tmp.alloc(N));
    RTX_TRY(rank.alloc(max_nodes));
    RTX_TRY(total.alloc(1));
    RTX_TRY(tasks_a.alloc(N));
    RTX_TRY(tasks_b.alloc(N));
    RTX_TRY(so.alloc(max_nodes));
    RTX_TRY(bnodes.alloc(max_nodes));
    RTX_TRY(st.alloc(1));
    RTX_TRY(hipMemsetAsync(st.p, 0, sizeof(BuildStats), s));
    DevBuf<uint32_t> partial, merged, lcount, rb;
    DevBuf<uint4> wsplit;
    if (N >= kWideMinPrims) {
        // chunks of a wide level: G * tasks <= N / kWideMinChunk; tasks <= N / kWideMinPrims
        const uint32_t nchunks = std::min(kWideBlocks, N / kWideMinChunk);
        const uint32_t nwide = std::min(kWideTasks, N / kWideMinPrims);
        RTX_TRY(partial.alloc((size_t)nchunks * kPartialWords));
        RTX_TRY(merged.alloc((size_t)nwide * kPartialWords));
        RTX_TRY(lcount.alloc(nchunks));
        RTX_TRY(wsplit.alloc(nwide));
    }
    RTX_TRY(rb.alloc(6));
    // tasks of a small-task level above kTinyTask primitives (at most N / (kTinyTask + 1))
    DevBuf<uint32_t> big, nbig;
    RTX_TRY(big.alloc(N / (kTinyTask + 1) + 1));
    RTX_TRY(nbig.alloc(1));

    hipLaunchKernelGGL(k_iota, dim3(blocks_for(N, 256)), dim3(256), 0, s, idx.p, N);
    hipLaunchKernelGGL(k_root_init, dim3(1), dim3(64), 0, s, rb.p);
    hipLaunchKernelGGL(k_root_bounds, dim3(std::min<uint32_t>(blocks_for(N, 256), kRootBlocks)), dim3(256), 0, s,
                       d_plo, d_phi, N, rb.p);
    hipLaunchKernelGGL(k_root_task, dim3(1), dim3(64), 0, s, rb.p, N, tasks_a.p);
    RTX_TRY(hipGetLastError());

    // exclusive scan of so[0, n).kind into rank, the sum into total
    DevBuf<uint32_t> tsum, toff;
    RTX_TRY(tsum.alloc((size_t)max_nodes / kScanTile + 2));
    RTX_TRY(toff.alloc((size_t)max_nodes / kScanTile + 2));
    auto scan_kind = [&](uint32_t n) {
        if (n <= 16 * kScanTile) {
            hipLaunchKernelGGL(k_scan_kind, dim3(1), dim3(kScanBlock), 0, s, so.p, n, rank.p, total.p);
            return;
        }
        const uint32_t nt = (n + kScanTile - 1) / kScanTile;
        hipLaunchKernelGGL(k_scan_tiles<KindOf>, dim3(nt), dim3(kScanBlock), 0, s, KindOf{so.p}, n, tsum.p);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(kScanBlock), 0, s, tsum.p, nt, toff.p, total.p);
        hipLaunchKernelGGL(k_scan_apply<KindOf>, dim3(nt), dim3(kScanBlock), 0, s, KindOf{so.p}, n, toff.p, rank.p);
    };
    // exclusive scan of in[0, n) into out, the sum into total (the same tiles)
    auto scan_u32 = [&](const uint32_t* in, uint32_t n, uint32_t* out) {
        if (n <= 16 * kScanTile) {
            hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(kScanBlock), 0, s, in, n, out, total.p);
            return;
        }
        const uint32_t nt = (n + kScanTile - 1) / kScanTile;
        hipLaunchKernelGGL(k_scan_tiles<U32Of>, dim3(nt), dim3(kScanBlock), 0, s, U32Of{in}, n, tsum.p);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(kScanBlock), 0, s, tsum.p, nt, toff.p, total.p);
        hipLaunchKernelGGL(k_scan_apply<U32Of>, dim3(nt), dim3(kScanBlock), 0, s, U32Of{in}, n, toff.p, out);
    };

    uint32_t ntasks = 1, nnodes = 1;
    BuildTask* cur = tasks_a.p;
    BuildTask* nxt = tasks_b.p;
    std::vector<uint32_t> level_first;   // binary level d: nodes level_first[d] .. level_first[d+1]-1
    while (ntasks > 0) {
        level_first.push_back(nnodes - ntasks);
        // a few big tasks: each spread over G chunks (kWideBlocks in all, of
        // kWideMinChunk primitives at least on average), their bins merged by
        // k_merge_wide; a wave per task once tasks average fewer than
        // kSmallTask primitives, and on levels of kTinyLevel tasks or more a
        // thread per task of at most kTinyTask primitives (the same tree every
        // way: binning, SAH and partition do not depend on the split)
        if (ntasks <= kWideTasks && (uint64_t)N >= (uint64_t)kWideMinPrims * ntasks) {
            const uint32_t G = std::min(kWideBlocks / ntasks, N / (ntasks * kWideMinChunk));
            const dim3 gw(G, ntasks);
            hipLaunchKernelGGL(k_bin_wide<kBuildBlock>, gw, dim3(kBuildBlock), 0, s, cur, d_plo, d_phi, idx.p, G,
                               partial.p);
            hipLaunchKernelGGL(k_merge_wide, dim3(kPartialWords / kMergeWords, ntasks), dim3(256), 0, s, G, partial.p,
                               merged.p);
            hipLaunchKernelGGL(k_decide_wide<kBuildBlock>, dim3(ntasks), dim3(kBuildBlock), 0, s, cur, 1u, merged.p,
                               bnodes.p, so.p, wsplit.p, max_leaf);
            hipLaunchKernelGGL(k_count_wide<kBuildBlock>, gw, dim3(kBuildBlock), 0, s, cur, d_plo, d_phi, idx.p, G,
                               wsplit.p, lcount.p);
            hipLaunchKernelGGL(k_scatter_wide<kBuildBlock>, gw, dim3(kBuildBlock), 0, s, cur, d_plo, d_phi, idx.p,
                               tmp.p, G, wsplit.p, lcount.p);
            hipLaunchKernelGGL(k_copy_wide, gw, dim3(256), 0, s, cur, G, wsplit.p, tmp.p, idx.p);
        } else if ((uint64_t)N < (uint64_t)kSmallTask * ntasks && ntasks < kTinyLevel) {
            hipLaunchKernelGGL(k_level<64>, dim3(ntasks), dim3(64), 0, s, cur, d_plo, d_phi, idx.p, tmp.p, bnodes.p,
                               so.p, max_leaf);
        } else if ((uint64_t)N < (uint64_t)kSmallTask * ntasks) {
            RTX_TRY(hipMemsetAsync(nbig.p, 0, sizeof(uint32_t), s));
            hipLaunchKernelGGL(k_level_tiny, dim3(blocks_for(ntasks, 256)), dim3(256), 0, s, cur, ntasks, d_plo, d_phi,
                               idx.p, tmp.p, bnodes.p, so.p, max_leaf, big.p, nbig.p);
            hipLaunchKernelGGL(k_level_list, dim3(std::min(ntasks, kListGrid)), dim3(64), 0, s, cur, big.p, nbig.p,
                               d_plo, d_phi, idx.p, tmp.p, bnodes.p, so.p, max_leaf);
        } else
            hipLaunchKernelGGL(k_level<kBuildBlock>, dim3(ntasks), dim3(kBuildBlock), 0, s, cur, d_plo, d_phi, idx.p,
                               tmp.p, bnodes.p, so.p, max_leaf);
        scan_kind(ntasks);
        hipLaunchKernelGGL(k_emit, dim3(blocks_for(ntasks, 256)), dim3(256), 0, s, cur, so.p, rank.p, ntasks,
                           nnodes, bnodes.p, nxt, st.p);
        RTX_TRY(hipGetLastError());
        uint32_t nsplit = 0;
        RTX_TRY(hipMemcpyAsync(&nsplit, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        RTX_TRY(hipStreamSynchronize(s));
        nnodes += 2 * nsplit;
        ntasks = 2 * nsplit;
        std::swap(cur, nxt);
    }

    hipLaunchKernelGGL(k_sah, dim3(std::min(blocks_for(nnodes, 256), kSahBlocks)), dim3(256), 0, s, bnodes.p, nnodes,
                       st.p);
    // compact inner nodes into the traversal layout (BFS order of the build)
    hipLaunchKernelGGL(k_inner_flag, dim3(blocks_for(nnodes, 256)), dim3(256), 0, s, bnodes.p, nnodes, so.p);
    scan_kind(nnodes);
    RTX_TRY(hipGetLastError());
    uint32_t ninner = 0;
    BuildStats hst;
    RTX_TRY(hipMemcpyAsync(&ninner, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RTX_TRY(hipMemcpyAsync(&hst, st.p, sizeof(BuildStats), hipMemcpyDeviceToHost, s));
    BNode root;
    RTX_TRY(hipMemcpyAsync(&root, bnodes.p, sizeof(BNode), hipMemcpyDeviceToHost, s));
    RTX_TRY(hipStreamSynchronize(s));

    DevBuf<float4> nodes, tris;
    if (opt.gather) {
        RTX_TRY(tris.alloc(3 * ((size_t)N + 1)));
        hipLaunchKernelGGL(k_gather_tris, dim3(blocks_for((uint64_t)N + 1, 256)), dim3(256), 0, s, d_tri_world,
                           idx.p, N, tris.p);
        RTX_TRY(hipGetLastError());
    }

    level_first.push_back(nnodes);
#if RTX_DP_COLLAPSE
    // the collapse's choices, bottom-up by binary level
    DevBuf<float4> dpc;
    DevBuf<uint32_t> dp_ntri, dp_first, dp_meta;
    if (ninner) {
        RTX_TRY(dpc.alloc(nnodes));
        RTX_TRY(dp_ntri.alloc(nnodes));
        RTX_TRY(dp_first.alloc(nnodes));
        RTX_TRY(dp_meta.alloc(nnodes));
        for (size_t d = level_first.size() - 1; d-- > 0;) {
            const uint32_t b = level_first[d], n = level_first[d + 1] - b;
            if (n == 0) continue;
            hipLaunchKernelGGL(k_dp_level, dim3(blocks_for(n, 256)), dim3(256), 0, s, bnodes.p, b, n, dpc.p,
                               dp_ntri.p, dp_first.p, dp_meta.p, max_leaf);
        }
        RTX_TRY(hipGetLastError());
    }
    const uint32_t *wmeta = dp_meta.p, *wntri = dp_ntri.p, *wfirst = dp_first.p;
#else
    const uint32_t *wmeta = nullptr, *wntri = nullptr, *wfirst = nullptr;
#endif

    // 4-wide nodes: at most one per binary inner node
    uint32_t nwide = 0, wide_depth = 0;
    if (ninner) {
        DevBuf<float4> wide;
        DevBuf<uint32_t> fa, fb, cnt, off;
        DevBuf<uint4> cid;
        RTX_TRY(wide.alloc(kWideNodeF4 * (size_t)ninner));
        RTX_TRY(fa.alloc(ninner));
        RTX_TRY(fb.alloc(ninner));
        RTX_TRY(cnt.alloc(ninner));
        RTX_TRY(off.alloc(ninner));
        RTX_TRY(cid.alloc(ninner));
        const uint32_t zero = 0;
        RTX_TRY(hipMemcpyAsync(fa.p, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, s));   // the root
        uint32_t n = 1;
        uint32_t* cur = fa.p;
        uint32_t* nxt = fb.p;
        while (n > 0) {
            out.level_begin.push_back(nwide);
#if RTX_DP_COLLAPSE
            hipLaunchKernelGGL(k_collapse4_dp, dim3(blocks_for(n, 256)), dim3(256), 0, s, bnodes.p, cur, n, dp_meta.p,
                               cid.p, cnt.p);
#else
            hipLaunchKernelGGL(k_collapse4, dim3(blocks_for(n, 256)), dim3(256), 0, s, bnodes.p, cur, n, cid.p,
                               cnt.p);
#endif
            scan_u32(cnt.p, n, off.p);
            hipLaunchKernelGGL(k_write4, dim3(blocks_for(n, 256)), dim3(256), 0, s, bnodes.p, cid.p, off.p, n,
                               nwide, nwide + n, leaf_ref(N, 1), wide.p, nxt, wmeta, wntri, wfirst);
            RTX_TRY(hipGetLastError());
            uint32_t nn = 0;
            RTX_TRY(hipMemcpyAsync(&nn, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            RTX_TRY(hipStreamSynchronize(s));
            nwide += n;
            wide_depth++;
            n = nn;
            std::swap(cur, nxt);
        }
        RTX_TRY(nodes.alloc(kWideNodeF4 * (size_t)nwide));
        RTX_TRY(hipMemcpyAsync(nodes.p, wide.p, sizeof(float4) * kWideNodeF4 * (size_t)nwide,
                               hipMemcpyDeviceToDevice, s));
    } else {
        RTX_TRY(nodes.alloc(kWideNodeF4));
    }
    RTX_TRY(hipStreamSynchronize(s));
    out.level_begin.push_back(nwide);

    const uint32_t root_cnt = __builtin_bit_cast(uint32_t, root.hi.w);
    const float rx = root.hi.x - root.lo.x, ry = root.hi.y - root.lo.y, rz = root.hi.z - root.lo.z;
    const double root_area = 2.0 * ((double)rx * ry + (double)ry * rz + (double)rz * rx);
    out.nodes = nodes.release();
    out.tris = tris.release();
    if (!opt.gather) out.order = idx.release();
    out.root_lo[0] = root.lo.x; out.root_lo[1] = root.lo.y; out.root_lo[2] = root.lo.z;
    out.root_hi[0] = root.hi.x; out.root_hi[1] = root.hi.y; out.root_hi[2] = root.hi.z;
    out.root = root_cnt ? leaf_ref(__builtin_bit_cast(uint32_t, root.lo.w), root_cnt) : 0u;
    out.num_nodes = ninner;
    out.num_wide = nwide;
    out.wide_depth = wide_depth;
    out.num_leaves = hst.leaves;
    out.max_depth = hst.max_depth;
    out.max_leaf = hst.max_leaf;
    out.sah = root_area > 0.0 ? hst.sah / root_area : 0.0;
    *outp = out;
    return hipSuccess;
}

}  // namespace rtx
