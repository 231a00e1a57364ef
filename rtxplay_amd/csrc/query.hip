// query.hip -- batch closest-hit query, device RNG draws and the post-process
// kernels.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "traverse.h"

namespace rtx {

namespace {

// Lanes of a wave that must finish a ray before the wave writes hits and
// takes new rays.  64 (the whole wave) measured fastest: with 32, 16 or 8
// the other lanes keep their traversal state, but every refill waits for
// the new rays' loads (1.15x, 1.6x and 2.2x slower on 2 M camera rays of
// the 7.9 M-triangle scene and their diffuse bounces).
#ifndef RTX_TRACE_BATCH
#define RTX_TRACE_BATCH 64
#endif

// closest hit for a batch of rays (the optixTrace query on its own):
// persistent grid, each lane takes rays from one work counter (one atomic
// per wave, lanes ranked by ballot); a wave traverses until `batch` of its
// lanes hold a final hit, writes those hits and gives those lanes new rays
// while the others keep their traversal state
template <int BLOCK, bool TL>
__global__ __launch_bounds__(BLOCK) void k_trace(const DevScene sc, const rtx_ray* __restrict__ rays, uint32_t n,
                                                 rtx_hit* __restrict__ hits, uint32_t* gstack, uint32_t gstride,
                                                 uint32_t* counter, int batch) {
    __shared__ uint32_t s_stack[kStackLDS * BLOCK];
    LaneStack<BLOCK> stk;
    stk.lds = (lds_u32*)(s_stack + threadIdx.x);
    stk.glob = gstack + blockIdx.x * BLOCK + threadIdx.x;
    stk.gstride = gstride;
    const uint32_t lane = threadIdx.x & 63u;
    Trav s;
    trav_begin(s, sc, f3(0.f, 0.f, 0.f), f3(0.f, 0.f, 1.f), 0.f, 0.f);
    bool going = false, held = false, exhausted = false;
    uint32_t i = 0;
    for (;;) {
        const bool need = !held && !exhausted;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (need) {
                i = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                if (i < n) {
                    const rtx_ray ray = rays[i];
                    trav_begin(s, sc, f3(ray.ori.x, ray.ori.y, ray.ori.z), f3(ray.dir.x, ray.dir.y, ray.dir.z),
                               ray.tmin, ray.tmax);
                    held = going = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(held)) break;
        trav_run<BLOCK, false, TL>(sc, s, going, held, batch, 32, stk);
        if (!held || going) continue;
        held = false;
        const HitFull h = s.h;
        rtx_hit o;
        if (h.tri == kNoHit) {
            o.t = __uint_as_float(0x7F800000u);
            o.u = o.v = 0.f;
            o.instance_id = o.primitive_index = o.prim = 0xFFFFFFFFu;
        } else {
            const float4* tp = sc.tris + 3 * (size_t)h.tri;
            o.t = h.t;
            o.u = h.u;
            o.v = h.v;
            o.instance_id = TL ? h.inst : __float_as_uint(tp[1].w);
            o.primitive_index = __float_as_uint(tp[2].w);
            // two-level: global id = the instance's first triangle + index in its mesh
            o.prim = TL ? __float_as_uint(sc.xinst[(size_t)kInstF4 * h.inst + 6].y) + h.prim : h.prim;
        }
        o.pad[0] = o.pad[1] = 0u;
        hits[i] = o;
    }
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

__global__ void k_rng(const uint32_t* __restrict__ seeds, uint32_t n, uint32_t k, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Frand48 r;
    r.init(seeds[i]);
    for (uint32_t j = 0; j < k; j++) out[(size_t)i * k + j] = r.next();
}

// optx/postproc.cu:2-16: IEC 61966-2-1 curve, then x255 truncated to uchar
__device__ __forceinline__ float srgb_curve(float x) {
    return x < .0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - .055f;
}
__device__ __forceinline__ unsigned char to_u8(float c) {
    const float v = c * 255.f;
    return (unsigned char)(v >= 255.f ? 255u : (v > 0.f ? (uint32_t)v : 0u));
}

// one thread per pixel, pix = x + w*y (optx/postproc.cu:36-47); the reference
// grid is 2D 32x32, this one 1D over the w*h pixels
template <bool SRGB>
__global__ void k_pp(const float* __restrict__ src, uchar4* __restrict__ dst, uint32_t npix) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float r = src[3 * (size_t)i], g = src[3 * (size_t)i + 1], b = src[3 * (size_t)i + 2];
    if (SRGB) dst[i] = make_uchar4(to_u8(srgb_curve(r)), to_u8(srgb_curve(g)), to_u8(srgb_curve(b)), 255);
    else dst[i] = make_uchar4(to_u8(r), to_u8(g), to_u8(b), 255);
}

}  // namespace

uint32_t trace_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace<kTraceBlock, false>, kTraceBlock, 0) != hipSuccess ||
        nb < 1)
        nb = 1;
    return (uint32_t)nb;
}

hipError_t launch_trace(const DevScene& sc, const rtx_ray* rays, uint32_t n, rtx_hit* hits, const LaunchScratch& ls,
                        hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint32_t grid = (n + kTraceBlock - 1) / kTraceBlock;
    if (grid > ls.grid) grid = ls.grid;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, ls.counter);
    if (sc.xinst)
        hipLaunchKernelGGL((k_trace<kTraceBlock, true>), dim3(grid), dim3(kTraceBlock), 0, s, sc, rays, n, hits,
                           ls.gstack, ls.gstride, ls.counter, (int)RTX_TRACE_BATCH);
    else
        hipLaunchKernelGGL((k_trace<kTraceBlock, false>), dim3(grid), dim3(kTraceBlock), 0, s, sc, rays, n, hits,
                           ls.gstack, ls.gstride, ls.counter, (int)RTX_TRACE_BATCH);
    return hipGetLastError();
}

hipError_t launch_rng(const uint32_t* seeds, uint32_t n, uint32_t k, float* out, hipStream_t s) {
    if (n == 0 || k == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rng, dim3((n + 255) / 256), dim3(256), 0, s, seeds, n, k, out);
    return hipGetLastError();
}

hipError_t launch_pp(const rtx_float3* src, rtx_uchar4* dst, int w, int h, bool srgb, hipStream_t s) {
    const uint32_t npix = (uint32_t)w * (uint32_t)h;
    if (npix == 0) return hipSuccess;
    const dim3 grid((npix + 255) / 256), block(256);
    if (srgb)
        hipLaunchKernelGGL(k_pp<true>, grid, block, 0, s, (const float*)src, (uchar4*)dst, npix);
    else
        hipLaunchKernelGGL(k_pp<false>, grid, block, 0, s, (const float*)src, (uchar4*)dst, npix);
    return hipGetLastError();
}

}  // namespace rtx
