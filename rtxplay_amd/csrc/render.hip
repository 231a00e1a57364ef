// render.hip -- one launch renders a frame: raygen + closest hit + scatter +
// spp accumulation, the work of __raygen__camera (optx/camera_r.cu:23-121),
// __miss__ambient (optx/camera_r.cu:123-147) and the three closest-hit
// programs (optx/optics_r.cu:22-431) under one optixLaunch.
//
// Shape: a persistent kernel.  Each lane owns one pixel at a time and walks
// its samples one path after the other, so the pixel's Frand48 stream
// (seeded with the pixel index) is consumed in exactly the reference order.
// When a path ends the lane starts the pixel's next sample in the same loop
// iteration; when the pixel is done it takes the next pixel from a global
// counter (one atomic per wave, lanes ranked by ballot).  Pixels are handed
// out in 8x8 tiles so that the primary rays of a wave are coherent.  Every
// wave loops until the counter passes the last tile, then exits.
//
// The recursion of the reference becomes a loop with a throughput product;
// the reference multiplies albedos on the way back up the recursion, this
// multiplies them on the way down (same factors, other association).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "traverse.h"

namespace rtx {

namespace {

struct RenderArgs {
    DevScene sc;
    rtx_lp_general lp;
    float3 ambient;
    uint32_t* counter;
    uint32_t* gstack;
    uint32_t gstride;
    uint32_t tiles_x, tiles_y;
    uint32_t batch;       // lanes that must finish a ray before a wave shades (64: whole wave)
    uint32_t inner_frac;  // node phase ends below inner_frac/64 of the going lanes
    TraceCounters tc;
};

__device__ __forceinline__ float3 ld(const rtx_float3& v) { return f3(v.x, v.y, v.z); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return lo > x ? lo : x > hi ? hi : x; }
__device__ __forceinline__ void st(rtx_float3* p, float3 v) {
    p->x = v.x;
    p->y = v.y;
    p->z = v.z;
}

__global__ void k_zero(uint32_t* p) { *p = 0u; }

template <int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_render(const RenderArgs a) {
    __shared__ uint32_t s_stack[kStackLDS * BLOCK];
    LaneStack<BLOCK> stk;
    stk.lds = (lds_u32*)(s_stack + threadIdx.x);
    stk.glob = a.gstack + blockIdx.x * BLOCK + threadIdx.x;
    stk.gstride = a.gstride;

    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t W = a.lp.image_w, H = a.lp.image_h;
    const uint32_t nwork = a.tiles_x * a.tiles_y * 64u;
    const uint32_t spp = a.lp.spp, maxd = a.lp.depth;
    const float inv_spp = 1.f / (float)spp;   // optx/camera_r.cu:112 with optx/v.h:44 (1/b * a)
    const float3 eye = ld(a.lp.camera.eye), cu = ld(a.lp.camera.u), cv = ld(a.lp.camera.v);
    const float3 hvec = ld(a.lp.camera.hvec), wvec = ld(a.lp.camera.wvec), dvec = ld(a.lp.camera.dvec);
    const float half_ap = a.lp.camera.aperture / 2.f;

    bool active = false, exhausted = false, going = false;
    uint32_t x = 0, y = 0, sample = 0, k = 0, rpp = 0;
    uint32_t cnt_nodes = 0, cnt_tris = 0;
    Frand48 rng;
    rng.state = 0;
    float3 thr = f3(1, 1, 1);
    float3 acc = f3(0, 0, 0), nacc = f3(0, 0, 0), aacc = f3(0, 0, 0), gN = f3(0, 0, 0), gA = f3(0, 0, 0);
    Trav tv;
    trav_begin(tv, a.sc, f3(0, 0, 0), f3(0, 0, 1), 1e-3f, 1e16f);
    uint64_t cyc_trav = 0;
    uint32_t wave_iters[2] = {0, 0};
    const uint64_t cyc_start = COUNT ? clock64() : 0;

    // optx/camera_r.cu:69-99 and optx/camera.h:67-77: the sample's primary ray
    auto start_sample = [&]() {
        const float s = 2.f * ((float)x + rng.next()) / (float)W - 1.f;
        const float t = 2.f * ((float)y + rng.next()) / (float)H - 1.f;
        const float3 rd = half_ap * rnd_in_disk(rng);
        const float3 off = rd.x * cu + rd.y * cv;
        trav_begin(tv, a.sc, eye + off, ((s * wvec + t * hvec) - dvec) - off, 1e-3f, 1e16f);
        going = true;
        k = 0;
        thr = f3(1.f, 1.f, 1.f);
        gN = f3(0.f, 0.f, 0.f);
        gA = f3(0.f, 0.f, 0.f);
    };

    for (;;) {
        // lanes without a pixel take one; out-of-image tile slots are skipped
        for (;;) {
            const bool need = !active && !exhausted;
            const uint64_t mask = __ballot(need);
            if (mask == 0) break;
            const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(a.counter, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (need) {
                const uint32_t wi = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                if (wi >= nwork) {
                    exhausted = true;
                } else {
                    const uint32_t tile = wi >> 6;
                    x = (tile % a.tiles_x) * 8u + (wi & 7u);
                    y = (tile / a.tiles_x) * 8u + ((wi >> 3) & 7u);
                    if (x < W && y < H) {
                        active = true;
                        rng.init(W * y + x);              // optx/camera_r.cu:35, :42-43
                        sample = 0;
                        rpp = 0;
                        cnt_nodes = cnt_tris = 0;
                        acc = nacc = aacc = f3(0.f, 0.f, 0.f);
                        start_sample();
                    }
                }
            }
        }
        if (!__any(active)) break;

        // traverse until a batch of lanes has a final closest hit; the others
        // keep their traversal state and resume after the shading below
        const uint64_t c0 = COUNT ? clock64() : 0;
        trav_run<BLOCK, COUNT>(a.sc, tv, going, active, (int)a.batch, (int)a.inner_frac, stk, &cnt_nodes, &cnt_tris,
                               wave_iters);
        if (COUNT) cyc_trav += clock64() - c0;
        if (!active || going) continue;

        k++;                                               // payload 5 (rays per trace)
        const float3 d = tv.r.d;
        bool end = false;
        float3 col = f3(0.f, 0.f, 0.f);
        if (tv.h.tri == kNoHit) {
            // __miss__ambient, optx/camera_r.cu:135-146
            const float3 unit = unitv(d);
            const float t = .5f * (unit.y + 1.f);
            col = thr * ((1.f - t) * f3(1.f, 1.f, 1.f) + t * a.ambient);
            end = true;
        } else if (k == maxd) {
            end = true;                                    // optx/optics_r.cu:34-42
        } else {
            // hit reconstruction, optx/optics_r.cu:45-90
            const float4* tp = a.sc.tris + 3 * (size_t)tv.h.tri;
            const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
            const uint32_t inst = __float_as_uint(t1.w);
            const float4 m0 = a.sc.insts[2 * inst], m1 = a.sc.insts[2 * inst + 1];
            const float3 A = xyz(t0), B = xyz(t1), C = xyz(t2);
            const float w = 1.f - tv.h.u - tv.h.v;
            const float3 hp = (w * A + tv.h.u * B) + tv.h.v * C;
            float3 N = unitv(cross(B - A, C - A));
            if (dot(d, N) > 0.f) N = -N;
            const int type = __float_as_int(m1.w);
            float3 nd;
            if (type == RTX_OPTICS_DIFFUSE) {
                // optx/optics_r.cu:108, :135-146
                nd = N + unitv(rnd_in_sphere(rng));
                const float3 alb = xyz(m0);
                thr = thr * alb;
                gN = N;
                gA = alb;
            } else if (type == RTX_OPTICS_REFLECT) {
                // optx/optics_r.cu:235-283
                const float3 d1 = unitv(d);
                const float3 refl = d1 - (2.f * dot(d1, N)) * N;
                nd = refl + m0.w * rnd_in_sphere(rng);
                const float3 alb = xyz(m0);
                gN = N;
                gA = alb;
                if (dot(nd, N) > 0.f) thr = thr * alb;
                else end = true;                           // absorbed: black
            } else {
                // optx/optics_r.cu:372-398
                const float3 d1 = unitv(d);
                const float cos_t = fminf(dot(-d1, N), 1.f);
                const float sin_t = sqrtf(1.f - cos_t * cos_t);
                const float3 O = hp - xyz(m1);             // center = instance translation
                const float ratio = 0.f > dot(d, O) ? 1.f / m0.w : m0.w;
                const bool cannot = ratio * sin_t > 1.f;
                float r0 = (1.f - ratio) / (1.f + ratio);
                r0 = r0 * r0;
                const float schlick = r0 + (1.f - r0) * powf(1.f - cos_t, 5.f);
                if (cannot || schlick > rng.next()) {
                    nd = d1 - (2.f * dot(d1, N)) * N;
                } else {
                    const float theta = fminf(dot(-d1, N), 1.f);
                    const float3 perpen = ratio * (d1 + theta * N);
                    const float3 parall = (-sqrtf(fabsf(1.f - dot(perpen, perpen)))) * N;
                    nd = perpen + parall;
                }
            }
            if (!end) {
                trav_begin(tv, a.sc, hp, nd, 1e-3f, 1e16f);   // optx/optics_r.cu:115-132
                going = true;
            }
        }

        if (end) {
            acc = acc + col;                               // optx/camera_r.cu:102-108
            rpp += k;
            nacc = nacc + gN;
            aacc = aacc + gA;
            if (++sample == spp) {
                const uint32_t pix = W * y + x;            // optx/camera_r.cu:111-120
                const float3 mean = inv_spp * acc;
                st(&a.lp.rawRGB[pix], f3(clampf(mean.x, 0.f, 1.f), clampf(mean.y, 0.f, 1.f), clampf(mean.z, 0.f, 1.f)));
                if (a.lp.rpp) a.lp.rpp[pix] = rpp;
                if (a.lp.normals) st(&a.lp.normals[pix], inv_spp * nacc);
                if (a.lp.albedos) st(&a.lp.albedos[pix], inv_spp * aacc);
                if (COUNT) {
                    a.tc.nodes[pix] = cnt_nodes;
                    a.tc.tris[pix] = cnt_tris;
                }
                active = false;
            } else {
                start_sample();
            }
        }
    }
    if (COUNT && a.tc.wave_cycles) {
        if (lane == 0) {
            atomicAdd((unsigned long long*)&a.tc.wave_cycles[0], (unsigned long long)cyc_trav);
            atomicAdd((unsigned long long*)&a.tc.wave_cycles[1], (unsigned long long)(clock64() - cyc_start));
        }
        atomicAdd((unsigned long long*)&a.tc.wave_cycles[2], (unsigned long long)wave_iters[0]);
        atomicAdd((unsigned long long*)&a.tc.wave_cycles[3], (unsigned long long)wave_iters[1]);
    }
}

// picker: one primary ray through (pick_x, pick_y), optx/camera_r.cu:26-32 and
// optx/optics_r.cu:24-28 / optx/camera_r.cu:125-129
__global__ __launch_bounds__(64) void k_pick(const RenderArgs a) {
    __shared__ uint32_t s_stack[kStackLDS * 64];
    if (threadIdx.x != 0) return;
    LaneStack<64> stk;
    stk.lds = (lds_u32*)s_stack;
    stk.glob = a.gstack;
    stk.gstride = a.gstride;
    const uint32_t W = a.lp.image_w, H = a.lp.image_h;
    const uint32_t x = a.lp.pick_x, y = a.lp.pick_y;
    Frand48 rng;
    rng.init(W * y + x);
    const float s = 2.f * ((float)x + rng.next()) / (float)W - 1.f;
    const float t = 2.f * ((float)y + rng.next()) / (float)H - 1.f;
    const float3 rd = (a.lp.camera.aperture / 2.f) * rnd_in_disk(rng);
    const float3 off = rd.x * ld(a.lp.camera.u) + rd.y * ld(a.lp.camera.v);
    const float3 o = ld(a.lp.camera.eye) + off;
    const float3 d = ((s * ld(a.lp.camera.wvec) + t * ld(a.lp.camera.hvec)) - ld(a.lp.camera.dvec)) - off;
    TRay r;
    tray_init(r, o, d, 1e-3f);
    HitFull h;
    traverse<64>(a.sc, r, 1e16f, h, stk);
    *a.lp.pick_id = h.tri == kNoHit ? 0xFFFFFFFFu : __float_as_uint(a.sc.tris[3 * (size_t)h.tri + 1].w);
}

RenderArgs make_args(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, const LaunchScratch& ls) {
    RenderArgs a;
    a.sc = sc;
    a.lp = lp;
    a.ambient = ambient;
    a.counter = ls.counter;
    a.gstack = ls.gstack;
    a.gstride = ls.gstride;
    a.tiles_x = (lp.image_w + 7u) / 8u;
    a.tiles_y = (lp.image_h + 7u) / 8u;
    a.batch = ls.shade_batch ? ls.shade_batch : 16u;
    a.inner_frac = ls.inner_frac ? ls.inner_frac : 36u;
    return a;
}

}  // namespace

uint32_t render_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kRenderBlock, false>, kRenderBlock, 0) !=
            hipSuccess || nb < 1)
        nb = 1;
    return (uint32_t)nb;
}

hipError_t launch_render(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, const LaunchScratch& ls,
                         hipStream_t s) {
    RenderArgs a = make_args(sc, lp, ambient, ls);
    a.tc = TraceCounters();
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    hipLaunchKernelGGL((k_render<kRenderBlock, false>), dim3(ls.grid), dim3(kRenderBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_render_counted(const DevScene& sc, const rtx_lp_general& lp, float3 ambient,
                                 const LaunchScratch& ls, const TraceCounters& tc, hipStream_t s) {
    RenderArgs a = make_args(sc, lp, ambient, ls);
    a.tc = tc;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    hipLaunchKernelGGL((k_render<kRenderBlock, true>), dim3(ls.grid), dim3(kRenderBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pick(const DevScene& sc, const rtx_lp_general& lp, const LaunchScratch& ls, hipStream_t s) {
    RenderArgs a = make_args(sc, lp, make_float3(0.f, 0.f, 0.f), ls);
    hipLaunchKernelGGL(k_pick, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace rtx
