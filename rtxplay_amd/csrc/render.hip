// render.hip -- one launch renders a frame: raygen + closest hit + scatter +
// spp accumulation, the work of __raygen__camera (optx/camera_r.cu:23-121),
// __miss__ambient (optx/camera_r.cu:123-147) and the three closest-hit
// programs (optx/optics_r.cu:22-431) under one optixLaunch.  The same kernel
// renders a batch of frames (many scenes and cameras) in one launch.
//
// Shape: a persistent kernel.  Each lane owns one pixel at a time and walks
// its samples one path after the other, so the pixel's Frand48 stream
// (seeded with the pixel index) is consumed in exactly the reference order.
// Work items are 64-pixel tiles (tile_w x 64/tile_w) of a job (one frame);
// lanes take them from one global counter (one atomic per wave, lanes ranked
// by ballot).  A batch lays its jobs out largest first, so the big frames
// start early and the small ones fill the tail.  Every wave loops until the
// counter passes the last work item, then exits.
//
// A wave traverses until a batch of its lanes hold a final closest hit,
// shades those lanes (scatter, next ray, next sample or next pixel) and
// resumes the others where they stopped (trav_run in traverse.h).
//
// The recursion of the reference becomes a loop with a throughput product;
// the reference multiplies albedos on the way back up the recursion, this
// multiplies them on the way down (same factors, other association).
#include <hip/hip_runtime.h>

#include <cstring>

#include "kernels.h"
#include "traverse.h"

// waves per SIMD the single-frame kernel is compiled for (register budget:
// 4 waves leave 128 VGPRs, enough for the 4-wide node step without spills)
#ifndef RTX_RENDER_WAVES
#define RTX_RENDER_WAVES 4
#endif
#ifndef RTX_BATCH_WAVES
#define RTX_BATCH_WAVES RTX_RENDER_WAVES
#endif

namespace rtx {

namespace {

struct RenderArgs {
    Job job;              // the frame of a single launch
    const Job* jobs;      // batch: job table (device), sorted by work_begin
    uint32_t njobs;
    uint32_t nwork;       // total work items x 64
    uint32_t* counter;
    uint32_t* gstack;
    uint32_t gstride;
    uint32_t batch;       // lanes that must finish a ray before a wave shades (64: whole wave)
    uint32_t inner_frac;  // node phase ends below inner_frac/64 of the going lanes
    uint32_t tile_log2w;  // tile width = 1 << tile_log2w, tile height = 64 / width
    TraceCounters tc;
    // adjoint launch only
    const rtx_float3* dldi;   // dL/d rawRGB per pixel
    float* dparams;           // float4 per instance: d albedo (rgb), d fuzz -- accumulated
    uint32_t* tape;           // per-lane path records, word w of slot k at tape[(k*kTapeWords+w)*tape_stride + lane]
    uint32_t tape_stride;     // lanes covered by the tape (grid x block)
    uint32_t lds_grad;        // 1: accumulate per block in LDS first (num_insts x 4 floats)
    // plain frames: the tile of each tile slot (nullptr: slot order)
    const uint32_t* tile_order;
    // probe launch: rays per tile (one sample per pixel, nothing else written)
    uint32_t* tile_cost;
    // 1: the iterative programs' semantics (optx/camera_i.cu, optx/optics_i.cu),
    // 0: the recursive ones (optx/camera_r.cu, optx/optics_r.cu)
    uint32_t iterative;
};

// path record of the adjoint: meta (instance | type << 24 | flags), the
// throughput before the vertex (3), the unit incoming direction (3), 1/|d|,
// the facing normal (3), the metal's sphere sample r (3) or the dielectric's
// ratio
constexpr uint32_t kTapeWords = kAdjointTapeWords;
constexpr uint32_t kTapeRefract = 1u << 26;   // dielectric: refracted (else reflected)
constexpr uint32_t kTapeCosClamp = 1u << 27;  // dielectric: cos theta clamped to 1

__device__ __forceinline__ float3 ld(const rtx_float3& v) { return f3(v.x, v.y, v.z); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return lo > x ? lo : x > hi ? hi : x; }
// util::isnull, optx/util_gpu.h:39
__device__ __forceinline__ bool isnull(float3 v) { return v.x == 0.f && v.y == 0.f && v.z == 0.f; }
__device__ __forceinline__ void st(rtx_float3* p, float3 v) {
    p->x = v.x;
    p->y = v.y;
    p->z = v.z;
}

__global__ void k_zero(uint32_t* p) { *p = 0u; }

// job of a global work index: the last job whose work_begin <= wi
__device__ __forceinline__ uint32_t find_job(const Job* jobs, uint32_t njobs, uint32_t wi) {
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].work_begin <= wi) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// the render kernels test parked leaves with the pairs spread over the wave
// (leaves_compact in traverse.h; measured on the 1080p level-6 frame: lanes
// busy in the triangle tests 24 % -> 61 % in a statistics launch, the
// 500-spp frame 3 % faster); 0 keeps the per-lane loop
#ifndef RTX_COMPACT_LEAVES
#define RTX_COMPACT_LEAVES 1
#endif
// the two-level frame kernels take the (P - o) * idir slab form: its six
// fewer registers keep them from spilling next to the compacted leaf phase
// (1080p / 64 spp: level-6 scene 98.7 -> 96.7 ms, mixed levels 114.1 ->
// 109.8 ms against the FMA form)
#ifndef RTX_TL_FMA
#define RTX_TL_FMA 0
#endif
constexpr bool kCompactLeaves = RTX_COMPACT_LEAVES != 0;

template <int BLOCK, bool COUNT, bool BATCH, int SLDS, bool ADJ = false, bool TL = false, bool FMA = true,
          bool PROBE = false, int TOPN = 0>
__device__ __forceinline__ void render_body(const RenderArgs& a, uint32_t* s_stack, float* s_grad = nullptr,
                                            const lds_f4* top = nullptr) {
    constexpr bool CL = kCompactLeaves;
    __shared__ uint32_t s_cl[CL ? (BLOCK / 64) * kLeafScratchWords : 1];
    lds_u32* cl = (lds_u32*)(s_cl + (CL ? (threadIdx.x >> 6) * kLeafScratchWords : 0u));
    LaneStack<BLOCK, SLDS> stk;
    stk.lds = (lds_u32*)(s_stack + threadIdx.x);
    stk.glob = a.gstack + blockIdx.x * BLOCK + threadIdx.x;
    stk.gstride = a.gstride;

    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tw = 1u << a.tile_log2w;

    // the lane's job: uniform for a single launch, per lane for a batch
    const Job* J = BATCH ? a.jobs : &a.job;
    DevScene sc = a.job.sc;
    uint32_t W = a.job.lp.image_w, H = a.job.lp.image_h, spp = PROBE ? 1u : a.job.lp.spp, maxd = a.job.lp.depth;
    uint32_t tslot = 0;   // probe: tile slot of the lane's pixel

    bool active = false, exhausted = false, going = false;
    uint32_t x = 0, y = 0, sample = 0, k = 0, rpp = 0;
    uint32_t cnt_nodes = 0, cnt_tris = 0;
    Frand48 rng;
    rng.state = 0;
    float3 thr = f3(1, 1, 1);
    float3 acc = f3(0, 0, 0), nacc = f3(0, 0, 0), aacc = f3(0, 0, 0), gN = f3(0, 0, 0), gA = f3(0, 0, 0);
    Trav tv;
    trav_begin(tv, sc, f3(0, 0, 0), f3(0, 0, 1), 1e-3f, 1e16f);
    uint64_t cyc_trav = 0;
    uint32_t wave_iters[2] = {0, 0};
    const uint64_t cyc_start = COUNT ? clock64() : 0;
    // adjoint: per-pixel dL/d(sample radiance), path vertex count, tape column
    float3 delta = f3(0.f, 0.f, 0.f);
    uint32_t kv = 0;
    uint32_t* tape = ADJ ? a.tape + blockIdx.x * BLOCK + threadIdx.x : nullptr;
    auto tput = [&](uint32_t slot, uint32_t w, uint32_t v) {
        tape[(size_t)(slot * kTapeWords + w) * a.tape_stride] = v;
    };
    auto tget = [&](uint32_t slot, uint32_t w) -> uint32_t {
        return tape[(size_t)(slot * kTapeWords + w) * a.tape_stride];
    };
    auto tgetf = [&](uint32_t slot, uint32_t w) -> float { return __uint_as_float(tget(slot, w)); };
    auto tget3 = [&](uint32_t slot, uint32_t w) -> float3 { return f3(tgetf(slot, w), tgetf(slot, w + 1), tgetf(slot, w + 2)); };
    auto tput3 = [&](uint32_t slot, uint32_t w, float3 v) {
        tput(slot, w, __float_as_uint(v.x));
        tput(slot, w + 1, __float_as_uint(v.y));
        tput(slot, w + 2, __float_as_uint(v.z));
    };
    auto grad = [&](uint32_t inst, uint32_t c, float v) {
        if (a.lds_grad) atomicAdd(&s_grad[4 * inst + c], v);
        else atomicAdd(&a.dparams[4 * inst + c], v);
    };

    // the next ray of a lane: a new sample's primary ray (newsample) or the
    // scattered ray of a path that goes on (cont, from ro along rdv); both
    // are set up at one place per iteration, so a wave pays for one setup
    bool newsample = false, cont = false;
    float3 ro = f3(0.f, 0.f, 0.f), rdv = f3(0.f, 0.f, 1.f);

    for (;;) {
        // lanes without a pixel take one; tile slots outside the frame are skipped
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
                if (wi >= a.nwork) {
                    exhausted = true;
                } else {
                    if (BATCH) {
                        J = a.jobs + find_job(a.jobs, a.njobs, wi);
                        sc = J->sc;
                        W = J->lp.image_w;
                        H = J->lp.image_h;
                        spp = J->lp.spp;
                        maxd = J->lp.depth;
                    }
                    const uint32_t li = wi - J->work_begin;
                    const uint32_t in = li & 63u;
                    const uint32_t tile = !BATCH && a.tile_order ? a.tile_order[li >> 6] : li >> 6;
                    if (PROBE) tslot = li >> 6;
                    x = (tile % J->tiles_x) * tw + (in & (tw - 1u));
                    y = (tile / J->tiles_x) * (64u / tw) + (in >> a.tile_log2w);
                    if (x < W && y < H) {
                        active = true;
                        if (ADJ) {
                            // d loss / d sample radiance: through the clamp of the
                            // mean (zero where it clamps) and the 1/spp of the mean
                            const uint32_t pix = W * y + x;
                            const float3 g = ld(a.dldi[pix]), im = ld(J->lp.rawRGB[pix]);
                            const float is = 1.f / (float)spp;
                            delta = f3(im.x > 0.f && im.x < 1.f ? g.x * is : 0.f, im.y > 0.f && im.y < 1.f ? g.y * is : 0.f,
                                       im.z > 0.f && im.z < 1.f ? g.z * is : 0.f);
                        }
                        rng.init(W * y + x);              // optx/camera_r.cu:35, :42-43
                        sample = 0;
                        rpp = 0;
                        cnt_nodes = cnt_tris = 0;
                        acc = nacc = aacc = f3(0.f, 0.f, 0.f);
                        newsample = true;
                    }
                }
            }
        }
        if (!__any(active)) break;

        if (newsample) {
            // optx/camera_r.cu:69-99 and optx/camera.h:67-77: the sample's primary ray
            const rtx_camera& cam = J->lp.camera;
            const float s = 2.f * ((float)x + rng.next()) / (float)W - 1.f;
            const float t = 2.f * ((float)y + rng.next()) / (float)H - 1.f;
            const float3 rd = (cam.aperture / 2.f) * rnd_in_disk(rng);
            const float3 off = rd.x * ld(cam.u) + rd.y * ld(cam.v);
            ro = ld(cam.eye) + off;
            rdv = ((s * ld(cam.wvec) + t * ld(cam.hvec)) - ld(cam.dvec)) - off;
            k = 0;
            kv = 0;
            thr = f3(1.f, 1.f, 1.f);
            gN = f3(0.f, 0.f, 0.f);
            gA = f3(0.f, 0.f, 0.f);
        }
        if (newsample || cont) {
            trav_begin(tv, sc, ro, rdv, 1e-3f, 1e16f);     // optx/optics_r.cu:115-132 for a scattered ray
            going = true;
        }
        newsample = cont = false;

        // traverse until a batch of lanes has a final closest hit; the others
        // keep their traversal state and resume after the shading below
        const uint64_t c0 = COUNT ? clock64() : 0;
        trav_run<BLOCK, COUNT, TL, FMA, BATCH, TOPN, CL>(sc, tv, going, active, (int)a.batch,
                                                                     (int)a.inner_frac, stk, &cnt_nodes, &cnt_tris,
                                                                     wave_iters, top, cl);
        if (COUNT) cyc_trav += clock64() - c0;
        if (!active || going) continue;

        k++;                                               // payload 5 (rays per trace)
        const float3 d = tv.r.d;
        bool end = false;
        float3 col = f3(0.f, 0.f, 0.f);
        // adjoint of a finished path whose radiance is (product of the
        // albedos) * S0: walk the path records backwards with S = the product
        // of the factors after the vertex (S0 included) and lam = dL/d(outgoing
        // direction), which flows back through metal and dielectric vertices
        // up to the last diffuse one (chain: the last vertex's outgoing
        // direction matters)
        auto backprop = [&](float3 S, float3 lam, bool chain) {
            for (int j = (int)kv - 1; j >= 0; --j) {
                const uint32_t meta = tget(j, 0);
                const uint32_t inst = meta & 0xFFFFFFu;
                const int type = (int)((meta >> 24) & 3u);
                if (type != RTX_OPTICS_REFRACT) {
                    const float3 P = tget3(j, 1);
                    const float3 alb = xyz(ld4<BATCH>(sc.insts + 2 * inst));
                    const float3 ga = delta * P * S;
                    grad(inst, 0, ga.x);
                    grad(inst, 1, ga.y);
                    grad(inst, 2, ga.z);
                    S = S * alb;
                    if (type == RTX_OPTICS_DIFFUSE) chain = false;
                }
                if (!chain) continue;
                const float3 u = tget3(j, 4);
                const float iln = tgetf(j, 7);
                const float3 N = tget3(j, 8);
                float3 lu;                                  // dL/d(unit incoming direction)
                if (type == RTX_OPTICS_REFLECT) {
                    grad(inst, 3, dot(lam, tget3(j, 11)));
                    lu = lam - (2.f * dot(lam, N)) * N;
                } else if (!(meta & kTapeRefract)) {
                    lu = lam - (2.f * dot(lam, N)) * N;
                } else {
                    const float ratio = tgetf(j, 11);
                    const float c = (meta & kTapeCosClamp) ? 1.f : -dot(u, N);
                    const float3 pp = ratio * (u + c * N);
                    const float q = 1.f - dot(pp, pp);
                    const float sq = sqrtf(fabsf(q));
                    const float3 lpp = lam + ((q < 0.f ? -1.f : 1.f) / sq * dot(N, lam)) * pp;
                    lu = (meta & kTapeCosClamp) ? ratio * lpp : ratio * (lpp - dot(N, lpp) * N);
                }
                lam = iln * (lu - dot(u, lu) * u);
            }
        };
        if (tv.h.tri == kNoHit) {
            // __miss__ambient, optx/camera_r.cu:135-146
            const float3 unit = unitv(d);
            const float t = .5f * (unit.y + 1.f);
            const float3 sky = (1.f - t) * f3(1.f, 1.f, 1.f) + t * J->ambient;
            col = thr * sky;
            end = true;
            if (ADJ) {
                // radiance = (prod of albedos) * sky(unit(d)); dL/d d through the sky gradient
                const float3 amb = J->ambient;
                const float gam = .5f * (delta.x * thr.x * (amb.x - 1.f) + delta.y * thr.y * (amb.y - 1.f) +
                                         delta.z * thr.z * (amb.z - 1.f));
                const float il = 1.f / sqrtf(dot(d, d));
                backprop(sky, (gam * il) * (f3(0.f, 1.f, 0.f) - unit.y * unit), true);
            }
        } else if (k == maxd && !a.iterative) {
            end = true;                                    // optx/optics_r.cu:34-42
        } else {
            // hit reconstruction, optx/optics_r.cu:45-90
            const float4* tp = sc.tris + 3 * (size_t)tv.h.tri;
            const float4 t0 = ld4<BATCH>(tp), t1 = ld4<BATCH>(tp + 1), t2 = ld4<BATCH>(tp + 2);
            const uint32_t inst = TL ? tv.h.inst : __float_as_uint(t1.w);
            const float4 m0 = ld4<BATCH>(sc.insts + 2 * inst), m1 = ld4<BATCH>(sc.insts + 2 * inst + 1);
            float3 A = xyz(t0), B = xyz(t1), C = xyz(t2);
            if (TL) {
                // object-space vertices to world space (optx/optics_r.cu:59-75)
                const float4* X = sc.xinst + (size_t)kInstF4 * inst;
                const float4 w0 = ld4<BATCH>(X + 3), w1 = ld4<BATCH>(X + 4), w2 = ld4<BATCH>(X + 5);
                A = f3(xf_point(w0, A), xf_point(w1, A), xf_point(w2, A));
                B = f3(xf_point(w0, B), xf_point(w1, B), xf_point(w2, B));
                C = f3(xf_point(w0, C), xf_point(w1, C), xf_point(w2, C));
            }
            const float w = 1.f - tv.h.u - tv.h.v;
            const float3 hp = (w * A + tv.h.u * B) + tv.h.v * C;
            float3 N = unitv(cross(B - A, C - A));
            if (dot(d, N) > 0.f) N = -N;
            const int type = __float_as_int(m1.w);
            float3 nd;
            uint32_t meta = inst | ((uint32_t)type << 24);
            if (ADJ) tput3(kv, 1, thr);
            // diffuse and metal lanes draw their point in the unit sphere in
            // one rejection loop, and one unit() serves the diffuse sample
            // and the incoming direction of metal and dielectric lanes (each
            // lane's draws keep the reference order)
            const bool diff = type == RTX_OPTICS_DIFFUSE, metal = type == RTX_OPTICS_REFLECT;
            float3 rs = f3(0.f, 0.f, 0.f);
            if (diff || metal) rs = rnd_in_sphere(rng);
            const float3 un = unitv(diff ? rs : d);
            const float3 refl = un - (2.f * dot(un, N)) * N;   // optx/optics_r.cu:236, :392
            if (diff || metal) {
                // denoiser guide values.  Recursive programs set them after
                // the recursion returns, where still null (optx/optics_r.cu:
                // 143-146, :280-283): the deepest diffuse or metal hit's
                // normal and the deepest non-zero albedo.  Iterative ones set
                // them on the way down (optx/optics_i.cu:98-101, :173-176):
                // the first hit's normal and the first non-zero albedo.
                const float3 alb = xyz(m0);
                if (a.iterative) {
                    if (isnull(gN)) gN = N;
                    if (isnull(gA)) gA = alb;
                } else {
                    gN = N;
                    if (!isnull(alb)) gA = alb;
                }
            }
            if (diff) {
                // optx/optics_r.cu:108, :135-138
                nd = N + un;
                thr = thr * xyz(m0);
            } else if (metal) {
                // optx/optics_r.cu:235-279
                nd = refl + m0.w * rs;
                if (ADJ) tput3(kv, 11, rs);
                if (dot(nd, N) > 0.f) thr = thr * xyz(m0);
                else end = true;                           // absorbed: black
            } else {
                // optx/optics_r.cu:372-398
                const float3 d1 = un;
                const float cos_t = fminf(dot(-d1, N), 1.f);
                const float sin_t = sqrtf(1.f - cos_t * cos_t);
                const float3 O = hp - xyz(m1);             // center = instance translation
                const float ratio = 0.f > dot(d, O) ? 1.f / m0.w : m0.w;
                const bool cannot = ratio * sin_t > 1.f;
                float r0 = (1.f - ratio) / (1.f + ratio);
                r0 = r0 * r0;
                const float x = 1.f - cos_t, x2 = x * x;     // (1 - cos)^5 as x^2 x^2 x
                const float schlick = r0 + (1.f - r0) * ((x2 * x2) * x);
                if (cannot || schlick > rng.next()) {
                    nd = refl;
                } else {
                    const float theta = fminf(dot(-d1, N), 1.f);
                    const float3 perpen = ratio * (d1 + theta * N);
                    const float3 parall = (-sqrtf(fabsf(1.f - dot(perpen, perpen)))) * N;
                    nd = perpen + parall;
                    if (ADJ) meta |= kTapeRefract | (dot(-d1, N) >= 1.f ? kTapeCosClamp : 0u);
                }
                if (ADJ) tput(kv, 11, __float_as_uint(ratio));
            }
            if (ADJ) {
                tput(kv, 0, meta);
                if (type != RTX_OPTICS_DIFFUSE) {
                    const float il = 1.f / sqrtf(dot(d, d));
                    tput3(kv, 4, il * d);
                    tput(kv, 7, __float_as_uint(il));
                    tput3(kv, 8, N);
                }
                kv++;
            }
            if (!end && k == maxd) {
                // iterative semantics at the last ray (optx/camera_i.cu:75-95):
                // the loop ends and the sample keeps the throughput as its colour
                end = true;
                col = thr;
                if (ADJ) backprop(f3(1.f, 1.f, 1.f), f3(0.f, 0.f, 0.f), false);
            } else if (!end) {
                ro = hp;
                rdv = nd;
                cont = true;
            }
        }

        if (end) {
            acc = acc + col;                               // optx/camera_r.cu:102-108
            rpp += k;
            nacc = nacc + gN;
            aacc = aacc + gA;
            if (++sample == spp && ADJ) {
                active = false;
            } else if (sample == spp && PROBE) {
                atomicAdd(&a.tile_cost[tslot], rpp);
                active = false;
            } else if (sample == spp) {
                const rtx_lp_general& lp = J->lp;
                const uint32_t pix = W * y + x;            // optx/camera_r.cu:111-120
                const float inv_spp = 1.f / (float)spp;    // optx/v.h:44: 1/b * a
                const float3 mean = inv_spp * acc;
                st(&lp.rawRGB[pix], f3(clampf(mean.x, 0.f, 1.f), clampf(mean.y, 0.f, 1.f), clampf(mean.z, 0.f, 1.f)));
                if (lp.rpp) lp.rpp[pix] = rpp;
                if (lp.normals) st(&lp.normals[pix], inv_spp * nacc);
                if (lp.albedos) st(&lp.albedos[pix], inv_spp * aacc);
                if (COUNT) {
                    a.tc.nodes[pix] = cnt_nodes;
                    a.tc.tris[pix] = cnt_tris;
                }
                active = false;
            } else {
                newsample = true;
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

// nodes of the top levels that each block of the frame kernels copies to LDS
// (0: none; the node array is breadth-first, so these are nodes 0..N-1).
// With the 32-entry LDS stack (32 KB per block) and the compacted leaf
// phase's scratch (4 KB per block), 32 nodes (3.5 KB: the top two 4-wide
// levels and 11 of the third) fill the CU's 160 KB at 4 blocks per CU.
// Measured: on the 1080p / 500 spp frame of the level-6 scene (2.3 M nodes),
// before the leaf phase was compacted, 21 nodes -0.5 %, 42 -0.4 %, 64 -1.1 %
// time, and with it 32 nodes -0.8 %; on scenes whose arrays fit the MALL
// (the level-4 scene, 0.15 M nodes: +2.5 %; the batch's small scenes: +4 %)
// and on short frames of the level-6 scene (1 spp: +3 %) the waves whose
// lanes split between LDS and global nodes issue both loads for too little
// latency saved.  So frames of at least kOrderMinSpp samples per pixel of
// scenes of at least kLdsTopMinNodes nodes take the staged kernels.
#ifndef RTX_LDS_TOP
#define RTX_LDS_TOP 32
#endif
constexpr int kLdsTop = RTX_LDS_TOP;
constexpr uint32_t kLdsTopMinNodes = 1u << 20;   // 128 MB of nodes

// copies the scene's first TOPN nodes (fewer if it has fewer) to s_top
template <int BLOCK, int TOPN>
__device__ __forceinline__ void stage_top(const DevScene& sc, float4* s_top) {
    const uint32_t n = (sc.num_nodes < (uint32_t)TOPN ? sc.num_nodes : (uint32_t)TOPN) * (uint32_t)kWideNodeF4;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) s_top[i] = sc.nodes[i];
    __syncthreads();
}

// single frame: one block of 256 lanes per SIMD wave slot
template <int BLOCK, bool COUNT, int SLDS, bool TL, int TOPN>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(RTX_RENDER_WAVES))) void k_render(const RenderArgs a) {
    __shared__ uint32_t s_stack[SLDS * BLOCK];
    __shared__ float4 s_top[TOPN > 0 ? TOPN * kWideNodeF4 : 1];
    if (TOPN > 0) stage_top<BLOCK, TOPN>(a.job.sc, s_top);
    render_body<BLOCK, COUNT, false, SLDS, false, TL, !TL || RTX_TL_FMA, false, TOPN>(a, s_stack, nullptr,
                                                                                   (const lds_f4*)s_top);
}

// the probe of a plain frame: one sample per pixel, rays counted per tile
template <int BLOCK, int SLDS, bool TL, int TOPN>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(RTX_RENDER_WAVES))) void k_probe(const RenderArgs a) {
    __shared__ uint32_t s_stack[SLDS * BLOCK];
    __shared__ float4 s_top[TOPN > 0 ? TOPN * kWideNodeF4 : 1];
    if (TOPN > 0) stage_top<BLOCK, TOPN>(a.job.sc, s_top);
    render_body<BLOCK, false, false, SLDS, false, TL, true, true, TOPN>(a, s_stack, nullptr, (const lds_f4*)s_top);
}

__global__ void k_zero_n(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The tiles in descending probe cost, stable: cost quantised to
// kOrderBuckets classes (class 0 the most rays), a counting sort in one
// block.  Each chunk of 1024 tiles ranks its tiles within their class by
// wave ballots (the lanes of a class found bit by bit) and per-wave class
// counts, so equal classes keep the tiles' order and the result is the same
// in every run.
constexpr uint32_t kOrderBuckets = 64;
constexpr uint32_t kOrderBlock = 1024;
__global__ __launch_bounds__(kOrderBlock) void k_tile_order(const uint32_t* cost, uint32_t n, uint32_t cmax,
                                                            uint32_t* order) {
    __shared__ uint32_t start[kOrderBuckets];                 // running start of each class
    __shared__ uint32_t wcnt[kOrderBlock / 64][kOrderBuckets];  // per wave: class counts, then offsets
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    auto cls = [&](uint32_t t) -> uint32_t {
        const uint32_t c = cost[t] < cmax ? cost[t] : cmax;
        return (kOrderBuckets - 1u) - (uint32_t)(((uint64_t)c * (kOrderBuckets - 1u)) / cmax);
    };
    if (tid < kOrderBuckets) start[tid] = 0u;
    __syncthreads();
    for (uint32_t t = tid; t < n; t += kOrderBlock) atomicAdd(&start[cls(t)], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t s = 0;
        for (uint32_t k = 0; k < kOrderBuckets; k++) {
            const uint32_t c = start[k];
            start[k] = s;
            s += c;
        }
    }
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kOrderBlock) {
        const uint32_t t = base + tid;
        const bool valid = t < n;
        const uint32_t b = valid ? cls(t) : kOrderBuckets;   // past the end: a class of its own
        uint64_t m = ~0ull;
        for (int bit = 0; bit < 7; bit++) {
            const uint64_t v = __ballot((b >> bit) & 1u);
            m &= ((b >> bit) & 1u) ? v : ~v;
        }
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        for (uint32_t i = tid; i < (kOrderBlock / 64) * kOrderBuckets; i += kOrderBlock) (&wcnt[0][0])[i] = 0u;
        __syncthreads();
        if (valid && (m >> lane) == 1ull) wcnt[wv][b] = rank + 1u;   // the class's highest lane
        __syncthreads();
        if (tid < kOrderBuckets) {
            uint32_t s = start[tid];
            for (uint32_t w = 0; w < kOrderBlock / 64; w++) {
                const uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = s;
                s += c;
            }
            start[tid] = s;
        }
        __syncthreads();
        if (valid) order[wcnt[wv][b] + rank] = t;
        __syncthreads();
    }
}

// batch: the same loop with a job per lane, whose job state takes
// registers: the slab test in its (P - o) * idir form keeps flat scenes at
// 4 waves per SIMD (7 % faster than the FMA form under that register
// limit, and 20 % faster than 3 waves); two-level ones take the registers
// of 3 waves
template <int BLOCK, int SLDS, bool TL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(RTX_BATCH_WAVES))) void k_render_batch(const RenderArgs a) {
    __shared__ uint32_t s_stack[SLDS * BLOCK];
    render_body<BLOCK, false, true, SLDS, false, false, false>(a, s_stack);
}
template <int BLOCK, int SLDS>
__global__ __launch_bounds__(BLOCK) void k_render_batch_tl(const RenderArgs a) {
    __shared__ uint32_t s_stack[SLDS * BLOCK];
    render_body<BLOCK, false, true, SLDS, false, true, false>(a, s_stack);
}

// adjoint: replays every sample path of a frame (same per-pixel Frand48
// streams, same traversal) and back-propagates dL/d rawRGB along each
// escaped path into per-instance d albedo and d fuzz; per-block LDS sums
// (when the instances fit) are flushed with one global atomic per entry
template <int BLOCK, int SLDS, bool TL>
__global__ __launch_bounds__(BLOCK) void k_adjoint(const RenderArgs a) {
    __shared__ uint32_t s_stack[SLDS * BLOCK];
    extern __shared__ float s_grad[];
    const uint32_t ng = 4u * a.job.sc.num_insts;
    if (a.lds_grad) {
        for (uint32_t i = threadIdx.x; i < ng; i += BLOCK) s_grad[i] = 0.f;
        __syncthreads();
    }
    render_body<BLOCK, false, false, SLDS, true, TL>(a, s_stack, s_grad);
    if (a.lds_grad) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < ng; i += BLOCK) {
            const float v = s_grad[i];
            if (v != 0.f) atomicAdd(&a.dparams[i], v);
        }
    }
}

__global__ void k_fill(float* p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// material parameters as float4 per instance: albedo (rgb) and fuzz; only
// diffuse (albedo) and metal (albedo, fuzz) instances take them
__global__ void k_set_materials(float4* insts, uint32_t n, const float4* __restrict__ params) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int type = __float_as_int(insts[2 * i + 1].w);
    const float4 p = params[i];
    float4 m = insts[2 * i];
    if (type == RTX_OPTICS_DIFFUSE) m = make_float4(p.x, p.y, p.z, m.w);
    else if (type == RTX_OPTICS_REFLECT) m = p;
    insts[2 * i] = m;
}

__global__ void k_get_materials(const float4* __restrict__ insts, uint32_t n, float4* params) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) params[i] = insts[2 * i];
}

// picker: one primary ray through (pick_x, pick_y), optx/camera_r.cu:26-32 and
// optx/optics_r.cu:24-28 / optx/camera_r.cu:125-129
template <bool TL>
__global__ __launch_bounds__(64) void k_pick(const RenderArgs a) {
    __shared__ uint32_t s_stack[kStackLDS * 64];
    if (threadIdx.x != 0) return;
    LaneStack<64> stk;
    stk.lds = (lds_u32*)s_stack;
    stk.glob = a.gstack;
    stk.gstride = a.gstride;
    const rtx_lp_general& lp = a.job.lp;
    const uint32_t W = lp.image_w, H = lp.image_h;
    const uint32_t x = lp.pick_x, y = lp.pick_y;
    Frand48 rng;
    rng.init(W * y + x);
    const float s = 2.f * ((float)x + rng.next()) / (float)W - 1.f;
    const float t = 2.f * ((float)y + rng.next()) / (float)H - 1.f;
    const float3 rd = (lp.camera.aperture / 2.f) * rnd_in_disk(rng);
    const float3 off = rd.x * ld(lp.camera.u) + rd.y * ld(lp.camera.v);
    const float3 o = ld(lp.camera.eye) + off;
    const float3 d = ((s * ld(lp.camera.wvec) + t * ld(lp.camera.hvec)) - ld(lp.camera.dvec)) - off;
    TRay r;
    tray_init(r, o, d, 1e-3f);
    HitFull h;
    traverse<64, TL>(a.job.sc, r, 1e16f, h, stk);
    *lp.pick_id = h.tri == kNoHit ? 0xFFFFFFFFu
                                  : (TL ? h.inst : __float_as_uint(a.job.sc.tris[3 * (size_t)h.tri + 1].w));
}

uint32_t log2u(uint32_t v) {
    uint32_t l = 0;
    while ((1u << (l + 1)) <= v) l++;
    return l;
}

// the frame kernel for the scene's kind (flat / two-level) and the LDS stack size
// the frame kernel for the scene's kind (flat / two-level), the LDS stack
// size and whether the top nodes go to LDS (large scenes, not the counted launch)
template <bool COUNT, int TOPN>
void launch_frame_t(const RenderArgs& a, const LaunchScratch& ls, hipStream_t s) {
    const dim3 g(ls.grid), b(kRenderBlock);
    if (a.job.sc.xinst) {
        if (ls.stack_lds == 16) hipLaunchKernelGGL((k_render<kRenderBlock, COUNT, 16, true, TOPN>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_render<kRenderBlock, COUNT, 32, true, TOPN>), g, b, 0, s, a);
    } else {
        if (ls.stack_lds == 16) hipLaunchKernelGGL((k_render<kRenderBlock, COUNT, 16, false, TOPN>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_render<kRenderBlock, COUNT, 32, false, TOPN>), g, b, 0, s, a);
    }
}

bool stage_top_nodes(const Job& j) {
    return kLdsTop > 0 && j.sc.num_nodes >= kLdsTopMinNodes && j.lp.spp >= kOrderMinSpp;
}

template <bool COUNT>
hipError_t launch_frame(const RenderArgs& a, const LaunchScratch& ls, hipStream_t s) {
    if (!COUNT && stage_top_nodes(a.job)) launch_frame_t<COUNT, kLdsTop>(a, ls, s);
    else launch_frame_t<COUNT, 0>(a, ls, s);
    return hipGetLastError();
}

RenderArgs base_args(const LaunchScratch& ls) {
    RenderArgs a;
    std::memset(&a, 0, sizeof(a));
    a.counter = ls.counter;
    a.gstack = ls.gstack;
    a.gstride = ls.gstride;
    a.batch = ls.shade_batch ? ls.shade_batch : 40u;
    a.inner_frac = ls.inner_frac ? ls.inner_frac : 36u;
    a.tile_log2w = log2u(ls.tile_w ? ls.tile_w : 8u);
    a.iterative = ls.iterative;
    return a;
}

}  // namespace

Job make_job(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, uint32_t tile_w, uint32_t work_begin) {
    Job j;
    std::memset(&j, 0, sizeof(j));
    j.sc = sc;
    j.lp = lp;
    j.ambient = ambient;
    const uint32_t th = 64u / tile_w;
    j.tiles_x = (lp.image_w + tile_w - 1u) / tile_w;
    j.tiles_y = (lp.image_h + th - 1u) / th;
    j.work_begin = work_begin;
    j.work_end = work_begin + j.tiles_x * j.tiles_y * 64u;
    return j;
}

uint32_t render_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render<kRenderBlock, false, kStackLDS, false, kLdsTop>, kRenderBlock, 0) !=
            hipSuccess || nb < 1)
        nb = 1;
    return (uint32_t)nb;
}

namespace {
// a frame of kOrderMinSpp or more spp (with the scratch in place): the probe
// launch (one sample per pixel, rays per tile) and the tile order, which the
// frame's launch then follows (a.tile_order)
void order_tiles(RenderArgs& a, const LaunchScratch& ls, hipStream_t s) {
    const uint32_t ntiles = a.job.tiles_x * a.job.tiles_y;
    if (!frame_ordered(a.job.lp, ls) || !ls.tile_cost || ls.tile_cap < ntiles) return;
    RenderArgs p = a;
    p.tile_cost = ls.tile_cost;
    p.tile_order = nullptr;
    p.dldi = nullptr;
    p.dparams = nullptr;
    p.tape = nullptr;
    hipLaunchKernelGGL(k_zero_n, dim3((ntiles + 255u) / 256u), dim3(256), 0, s, ls.tile_cost, ntiles);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    const dim3 g(ls.grid), b(kRenderBlock);
    if (stage_top_nodes(a.job)) {
        if (a.job.sc.xinst) {
            if (ls.stack_lds == 16) hipLaunchKernelGGL((k_probe<kRenderBlock, 16, true, kLdsTop>), g, b, 0, s, p);
            else hipLaunchKernelGGL((k_probe<kRenderBlock, 32, true, kLdsTop>), g, b, 0, s, p);
        } else {
            if (ls.stack_lds == 16) hipLaunchKernelGGL((k_probe<kRenderBlock, 16, false, kLdsTop>), g, b, 0, s, p);
            else hipLaunchKernelGGL((k_probe<kRenderBlock, 32, false, kLdsTop>), g, b, 0, s, p);
        }
    } else {
        if (a.job.sc.xinst) {
            if (ls.stack_lds == 16) hipLaunchKernelGGL((k_probe<kRenderBlock, 16, true, 0>), g, b, 0, s, p);
            else hipLaunchKernelGGL((k_probe<kRenderBlock, 32, true, 0>), g, b, 0, s, p);
        } else {
            if (ls.stack_lds == 16) hipLaunchKernelGGL((k_probe<kRenderBlock, 16, false, 0>), g, b, 0, s, p);
            else hipLaunchKernelGGL((k_probe<kRenderBlock, 32, false, 0>), g, b, 0, s, p);
        }
    }
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderBlock), 0, s, ls.tile_cost, ntiles, 64u * a.job.lp.depth,
                       ls.tile_order);
    a.tile_order = ls.tile_order;
}
}  // namespace

hipError_t launch_render(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, const LaunchScratch& ls,
                         hipStream_t s) {
    RenderArgs a = base_args(ls);
    a.job = make_job(sc, lp, ambient, 1u << a.tile_log2w, 0);
    a.njobs = 1;
    a.nwork = a.job.work_end;
    order_tiles(a, ls, s);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    return launch_frame<false>(a, ls, s);
}

hipError_t launch_render_counted(const DevScene& sc, const rtx_lp_general& lp, float3 ambient,
                                 const LaunchScratch& ls, const TraceCounters& tc, hipStream_t s) {
    RenderArgs a = base_args(ls);
    a.job = make_job(sc, lp, ambient, 1u << a.tile_log2w, 0);
    a.njobs = 1;
    a.nwork = a.job.work_end;
    a.tc = tc;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    return launch_frame<true>(a, ls, s);
}

hipError_t launch_render_batch(const Job* d_jobs, const Job& first, uint32_t njobs, uint32_t nwork,
                               const LaunchScratch& ls, uint32_t tile_w, hipStream_t s) {
    RenderArgs a = base_args(ls);
    a.tile_log2w = log2u(tile_w);
    a.job = first;
    a.jobs = d_jobs;
    a.njobs = njobs;
    a.nwork = nwork;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    const dim3 g(ls.grid), b(kRenderBlock);
    if (first.sc.xinst) {
        if (ls.stack_lds == 16) hipLaunchKernelGGL((k_render_batch_tl<kRenderBlock, 16>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_render_batch_tl<kRenderBlock, 32>), g, b, 0, s, a);
    } else {
        if (ls.stack_lds == 16) hipLaunchKernelGGL((k_render_batch<kRenderBlock, 16, false>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_render_batch<kRenderBlock, 32, false>), g, b, 0, s, a);
    }
    return hipGetLastError();
}


hipError_t launch_adjoint(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, const LaunchScratch& ls,
                          const rtx_float3* dldi, float* dparams, uint32_t* tape, uint32_t tape_stride, hipStream_t s) {
    RenderArgs a = base_args(ls);
    a.job = make_job(sc, lp, ambient, 1u << a.tile_log2w, 0);
    a.njobs = 1;
    a.nwork = a.job.work_end;
    a.dldi = dldi;
    a.dparams = dparams;
    a.tape = tape;
    a.tape_stride = tape_stride;
    a.lds_grad = sc.num_insts <= kAdjointLdsInsts ? 1u : 0u;
    const uint32_t ng = 4u * sc.num_insts;
    if (ng) hipLaunchKernelGGL(k_fill, dim3((ng + 255) / 256), dim3(256), 0, s, dparams, ng, 0.f);
    order_tiles(a, ls, s);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, ls.counter);
    const size_t shm = a.lds_grad ? sizeof(float) * ng : 0;
    const dim3 g(ls.grid), b(kRenderBlock);
    if (sc.xinst) {
        if (ls.stack_lds == 16) hipLaunchKernelGGL((k_adjoint<kRenderBlock, 16, true>), g, b, shm, s, a);
        else hipLaunchKernelGGL((k_adjoint<kRenderBlock, 32, true>), g, b, shm, s, a);
    } else {
        if (ls.stack_lds == 16) hipLaunchKernelGGL((k_adjoint<kRenderBlock, 16, false>), g, b, shm, s, a);
        else hipLaunchKernelGGL((k_adjoint<kRenderBlock, 32, false>), g, b, shm, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_set_materials(const DevScene& sc, const float* params, hipStream_t s) {
    if (!sc.num_insts) return hipSuccess;
    hipLaunchKernelGGL(k_set_materials, dim3((sc.num_insts + 255) / 256), dim3(256), 0, s, (float4*)sc.insts,
                       sc.num_insts, (const float4*)params);
    return hipGetLastError();
}

hipError_t launch_get_materials(const DevScene& sc, float* params, hipStream_t s) {
    if (!sc.num_insts) return hipSuccess;
    hipLaunchKernelGGL(k_get_materials, dim3((sc.num_insts + 255) / 256), dim3(256), 0, s, sc.insts, sc.num_insts,
                       (float4*)params);
    return hipGetLastError();
}

hipError_t launch_pick(const DevScene& sc, const rtx_lp_general& lp, const LaunchScratch& ls, hipStream_t s) {
    RenderArgs a = base_args(ls);
    a.job = make_job(sc, lp, make_float3(0.f, 0.f, 0.f), 8u, 0);
    if (sc.xinst) hipLaunchKernelGGL(k_pick<true>, dim3(1), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_pick<false>, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace rtx
